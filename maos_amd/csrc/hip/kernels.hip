// Kernels of the tomography PCG/FDPCG solve for gfx950.
//
// Reference counterparts (each replaced by a gather-only kernel):
//   gpu_map2map_do 'n'/'t'   cuda/prop_wrap.cu:34-349   -> k_wfs (HX), k_layer (HX')
//   gpu_gp_do / gpu_gpt_do   cuda/tomo.cu:361-573        -> k_wfs, k_ttr
//   gpu_laplacian_do         cuda/tomo.cu:317-355        -> k_layer
//   cuFFT R2C/C2R +
//   fdpcg_mul_block_sync_half cuda/fdpcg.cu:112-226      -> k_row, k_col, k_irow
//   inn_do / add_do / div_do cuda/kernel.cu, cuda/pcg.cu -> FP64 partials in
//                                                          k_layer/k_irow/k_row,
//                                                          fixed-order reduction
#include "kernels.hpp"
#include <type_traits>

#include "launch.hpp"

namespace mtk {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
template <typename C> struct CT;
template <> struct CT<float2> {
    typedef float R;
    static __device__ __forceinline__ float2 make(float a, float b) { return make_float2(a, b); }
};
template <> struct CT<double2> {
    typedef double R;
    static __device__ __forceinline__ double2 make(double a, double b) { return make_double2(a, b); }
};

// Job and block of this workgroup (see Job / WGI in kernels.hpp).
// The workgroup's table entry is read as one 8-byte scalar load (entries
// are 8-byte aligned: {job, bx} in the low word, {by, bz} in the high one).
#define KJOB                                                   \
    const unsigned long long wi_ = ((const CONSTAS unsigned long long *)wgi)[blockIdx.x]; \
    const Job J = jobs[wi_ & 0xffffu];                         \
    if (!J.on) return;                                         \
    CPD &PD = *(const CONSTAS PlanDev *)J.pd;                  \
    CG &G = PD.G;                                              \
    const int bx = (int)((wi_ >> 16) & 0xffffu), by = (int)((wi_ >> 32) & 0xffffu), bz = (int)(wi_ >> 48); \
    (void)bx;                                                  \
    (void)by;                                                  \
    (void)bz
#define KARGS const Job *__restrict__ jobs, const WGI *__restrict__ wgi

__device__ __forceinline__ float *vsel(CPD &P, int sel) { return sel == BUF_X ? P.x : sel == BUF_P ? P.p : P.z; }
template <typename RT>
__device__ __forceinline__ RT *rsel(CPD &P, int sel) {
    return (RT *)(sel == BUF_AP ? P.Ap : P.b);
}
__device__ __forceinline__ double *psel(CPD &P, int sel) {
    return sel == PART_PAP ? P.SC.part_pap : sel == PART_BB ? P.SC.part_bb : nullptr;
}
// Plain copies of small structs read through the constant address space.
__device__ __forceinline__ MRay mray(CG &G, int i) {
    const CONSTAS MRay &r = ((const CONSTAS MRay *)G.mray)[i];
    MRay m;
    m.cx = r.cx;
    m.cy = r.cy;
    m.ax0 = r.ax0;
    m.ax1 = r.ax1;
    m.ay0 = r.ay0;
    m.ay1 = r.ay1;
    m.wx = r.wx;
    m.wy = r.wy;
    m.fx = r.fx;
    m.fy = r.fy;
    m.nrow = r.nrow;
    m.qoff = r.qoff;
    return m;
}
__device__ __forceinline__ FDiv ldf(const CONSTAS FDiv &f) { return FDiv{f.d, f.m}; }
__device__ __forceinline__ FStage ldst(const CONSTAS FStage &f) {
    FStage st;
    st.R = f.R;
    st.Ns = f.Ns;
    st.nb = f.nb;
    st.tstep = f.tstep;
    st.dnb = ldf(f.dnb);
    st.dNs = ldf(f.dNs);
    return st;
}
__device__ __forceinline__ Scal ldsc(const CONSTAS Scal &c) {
    Scal r;
    r.part_pap = c.part_pap;
    r.part_rz = c.part_rz;
    r.part_rr = c.part_rr;
    r.part_bb = c.part_bb;
    r.rz = c.rz;
    r.alpha = c.alpha;
    r.res = c.res;
    r.tt = c.tt;
    r.maxpart = c.maxpart;
    r.maxit = c.maxit;
    r.npart_layer = c.npart_layer;
    r.npart_row = c.npart_row;
    r.npart_col = c.npart_col;
    return r;
}

// Buffer loads: a 32-bit byte offset against a base held in scalar
// registers (no per-lane 64-bit address arithmetic), and an offset at or
// past the buffer's size reads 0 (BOOB marks a ray that is not valid, so
// its term vanishes without a per-lane weight select).
constexpr unsigned BOOB = 0x80000000u;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t brsrc(const void *base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}
template <typename T> __device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, unsigned off);
template <> __device__ __forceinline__ double bload<double>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
template <> __device__ __forceinline__ float bload<float>(__amdgpu_buffer_rsrc_t r, unsigned off) {
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}

// ---------------------------------------------------------------------------
// Deterministic FP64 reductions (fixed order: per-lane strided sums, then a
// 64-lane shuffle tree, then wavefront partials summed by lane 0 in order).
// ---------------------------------------------------------------------------
__device__ double block_sum(double v, double *sh, int f32 = 0) {
    if (f32) {
        // precision mode 3: the reduction itself in FP32 (same order)
        float f = (float)v;
        for (int o = 32; o > 0; o >>= 1) f += __shfl_down(f, o, 64);
        v = f;
    } else {
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0;
        float sf = 0.f;
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; i++) {
            s += sh[i];
            sf += (float)sh[i];
        }
        sh[0] = f32 ? (double)sf : s;
    }
    __syncthreads();
    double r = sh[0];
    __syncthreads();
    return r;
}
__device__ double reduce_partials(const double *part, int n, double *sh, int f32 = 0) {
    double v = 0;
    float vf = 0.f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        v += part[i];
        vf += (float)part[i];
    }
    return block_sum(f32 ? (double)vf : v, sh, f32);
}
// Product / fused accumulation of a dot-product term, in FP32 for mode 3.
__device__ __forceinline__ double dmul(double a, double b, int f32) {
    return f32 ? (double)((float)a * (float)b) : a * b;
}
__device__ __forceinline__ double dfma(double a, double b, double c, int f32) {
    return f32 ? (double)fmaf((float)a, (float)b, (float)c) : a * b + c;
}
__device__ __forceinline__ double ddiv(double a, double b, int f32) {
    return f32 ? (double)((float)a / (float)b) : a / b;
}

// ---------------------------------------------------------------------------
// Stockham mixed-radix FFT on LDS: nf transforms of length n, stored
// contiguously; radices 2,3,4,5,7 (sizes that nextfftsize produces).
// tw[k] = exp(-2 pi i k/n).  SIGN=-1 forward, +1 inverse, unnormalized.
// ---------------------------------------------------------------------------
// cos/sin(2 pi k / R), k = 0..R-1, for the odd radices (the R-th roots of
// unity used inside one radix-R butterfly).
template <int R> struct Roots;
template <> struct Roots<3> {
    static constexpr double c[3] = {1.0, -0.5, -0.5};
    static constexpr double s[3] = {0.0, 0.86602540378443865, -0.86602540378443865};
};
template <> struct Roots<5> {
    static constexpr double c[5] = {1.0, 0.30901699437494742, -0.80901699437494742, -0.80901699437494742,
                                    0.30901699437494742};
    static constexpr double s[5] = {0.0, 0.95105651629515357, 0.58778525229247313, -0.58778525229247313,
                                    -0.95105651629515357};
};
template <> struct Roots<7> {
    static constexpr double c[7] = {1.0, 0.62348980185873353, -0.22252093395631440, -0.90096886790241913,
                                    -0.90096886790241913, -0.22252093395631440, 0.62348980185873353};
    static constexpr double s[7] = {0.0, 0.78183148246802981, 0.97492791218182361, 0.43388373911755812,
                                    -0.43388373911755812, -0.97492791218182361, -0.78183148246802981};
};

template <int R, int SIGN, typename C>
__device__ __forceinline__ void small_dft(C *v) {
    if constexpr (R == 2) {
        C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        C a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        C a2 = cadd(v[1], v[3]), d = csub(v[1], v[3]);
        // forward: a3 = -i*d ; inverse: +i*d
        C a3 = SIGN < 0 ? CT<C>::make(d.y, -d.x) : CT<C>::make(-d.y, d.x);
        v[0] = cadd(a0, a2);
        v[2] = csub(a0, a2);
        v[1] = cadd(a1, a3);
        v[3] = csub(a1, a3);
    } else {
        // odd R: pair v_r with v_(R-r); with W^(kr) = c + i s (s carrying
        // SIGN), W^(kr) v_r + W^(-kr) v_(R-r) = c (v_r + v_(R-r)) + i s (v_r - v_(R-r))
        static_assert(R % 2 == 1, "odd radix");
        typedef typename CT<C>::R F;
        constexpr int H = (R - 1) / 2;
        C a[H], d[H];
#pragma unroll
        for (int r = 1; r <= H; r++) {
            a[r - 1] = cadd(v[r], v[R - r]);
            d[r - 1] = csub(v[r], v[R - r]);
        }
        C o[R];
        o[0] = v[0];
#pragma unroll
        for (int r = 0; r < H; r++) o[0] = cadd(o[0], a[r]);
#pragma unroll
        for (int k = 1; k < R; k++) {
            F re = v[0].x, im = v[0].y;
#pragma unroll
            for (int r = 1; r <= H; r++) {
                const F c = (F)Roots<R>::c[(k * r) % R], sn = (F)(SIGN * Roots<R>::s[(k * r) % R]);
                re += c * a[r - 1].x - sn * d[r - 1].y;
                im += c * a[r - 1].y + sn * d[r - 1].x;
            }
            o[k] = CT<C>::make(re, im);
        }
#pragma unroll
        for (int k = 0; k < R; k++) v[k] = o[k];
    }
}

__device__ __forceinline__ unsigned fdiv(unsigned t, FDiv d) { return d.m ? __umulhi(t, d.m) : t; }

template <int R, int SIGN, typename C>
__device__ void fft_stage(const C *in, C *out, int n, int nf, const FStage &st, const C *tw) {
    const int nb = st.nb, Ns = st.Ns, tstep = st.tstep;
    const int tot = nf * nb;
    for (int t = threadIdx.x; t < tot; t += blockDim.x) {
        const int f = fdiv(t, st.dnb), j = t - f * nb;
        const C *src = in + f * n;
        C *dst = out + f * n;
        C v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = src[j + r * nb];
        const int jm = j - fdiv(j, st.dNs) * Ns;
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                C w = tw[jm * r * tstep];
                if (SIGN > 0) w.y = -w.y;
                v[r] = cmul(v[r], w);
            }
        }
        small_dft<R, SIGN, C>(v);
        const int idxD = (j - jm) * R + jm;
#pragma unroll
        for (int r = 0; r < R; r++) dst[idxD + r * Ns] = v[r];
    }
}

// Returns the buffer holding the result (a or b).  tw: the n twiddles, in LDS.
// RM: largest radix compiled in (3: radices 2, 3, 4; 7: also 5 and 7, whose
// butterflies need about twice the registers).
template <int SIGN, typename C, int RM>
__device__ C *fft_run(C *a, C *b, int n, int nf, int nst, const CONSTAS FStage *stages, const C *tw) {
    C *in = a, *out = b;
    for (int s = 0; s < nst; s++) {
        const FStage st = ldst(stages[s]);
        switch (st.R) {
            case 2: fft_stage<2, SIGN, C>(in, out, n, nf, st, tw); break;
            case 3: fft_stage<3, SIGN, C>(in, out, n, nf, st, tw); break;
            case 4: fft_stage<4, SIGN, C>(in, out, n, nf, st, tw); break;
            case 5:
                if constexpr (RM >= 5) fft_stage<5, SIGN, C>(in, out, n, nf, st, tw);
                break;
            case 7:
                if constexpr (RM >= 7) fft_stage<7, SIGN, C>(in, out, n, nf, st, tw);
                break;
            default: break;
        }
        __syncthreads();
        C *t = in;
        in = out;
        out = t;
    }
    return in;
}

// ---------------------------------------------------------------------------
// FFT of a column in registers: n = L * P points on L consecutive lanes (L a
// power of two dividing 64), lane l holding x[l + L p] in register p.
// Four-step form, with W_N = exp(-2 pi i / N):
//   X[k1 + P k2] = sum_l W_L^(l k2) W_n^(l k1) sum_p x[l + L p] W_P^(p k1)
// the forward transform: a radix-P DFT in each lane, the twiddles
// W_n^(l k1), then log2(L) radix-2 stages across the lanes by xor shuffles
// (decimation in frequency: no LDS, no barrier); on return lane l, register
// k1 holds X[k1 + P * bitrev_L(l)].  The inverse takes that layout back: the
// same stages in reverse order with conjugate twiddles (decimation in time),
// the conjugate twiddles, then the radix-P inverse DFT, leaving n x the
// inverse transform in natural order (lane l, register p: x[l + L p]).
// tw: the n twiddles exp(-2 pi i k / n) (the twiddle of a stage is the same
// on both lanes of a butterfly).
// ---------------------------------------------------------------------------
template <typename C> __device__ __forceinline__ C shfl_xor_c(C v, int m) {
    return CT<C>::make(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}
template <typename C> __device__ __forceinline__ C conj_c(C v) { return CT<C>::make(v.x, -v.y); }
// Exchange with lane ^ H for a compile-time H below 32, per 32-bit word:
// DPP quad permutes for H = 1, 2 (a VALU move), ds_swizzle in bit-mask mode
// (lane ^ H within each 32 lanes; no address register) for H = 4 .. 16.
template <int H> __device__ __forceinline__ int xchg_i(int v) {
    static_assert(H >= 1 && H < 32 && (H & (H - 1)) == 0, "power of two below 32");
    if constexpr (H == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
    else if constexpr (H == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
    else return __builtin_amdgcn_ds_swizzle(v, (H << 10) | 0x1F);
}
template <int H> __device__ __forceinline__ float xchg(float v) { return __int_as_float(xchg_i<H>(__float_as_int(v))); }
template <int H> __device__ __forceinline__ double xchg(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = xchg_i<H>((int)(b & 0xffffffffLL)), hi = xchg_i<H>((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <int H, typename C> __device__ __forceinline__ C xchg_c(C v) { return CT<C>::make(xchg<H>(v.x), xchg<H>(v.y)); }
// A butterfly is written without a select between its two results: the
// lane adds s * (own value) to its partner's value (s = +1 on the lower lane
// of the pair, -1 on the upper, so the pair forms a + b and a - b exactly)
// and multiplies by w' (the stage twiddle on the lane that takes it, else
// 1, an exact product).  q tracks L / (2h) by shifts.
template <int P, typename C>
__device__ __forceinline__ void fft_regs_fwd(C *y, int L, int l, const C *tw) {
    typedef typename CT<C>::R F;
    if constexpr (P > 1) small_dft<P, -1, C>(y);
#pragma unroll
    for (int k1 = 1; k1 < P; k1++) y[k1] = cmul(y[k1], tw[l * k1]);
    for (int h = L >> 1, q = 1; h >= 1; h >>= 1, q <<= 1) {
        const bool up = (l & h) != 0;
        const C wt = tw[(l & (h - 1)) * q * P];
        const C w = up ? wt : CT<C>::make(1, 0);
        const F sg = up ? F(-1) : F(1);
#pragma unroll
        for (int p = 0; p < P; p++) {
            const C o = shfl_xor_c(y[p], h);
            y[p] = cmul(CT<C>::make(o.x + sg * y[p].x, o.y + sg * y[p].y), w);
        }
    }
}
template <int P, typename C>
__device__ __forceinline__ void fft_regs_inv(C *y, int L, int l, const C *tw) {
    typedef typename CT<C>::R F;
    for (int h = 1, q = L >> 1; h < L; h <<= 1, q >>= 1) {
        const bool up = (l & h) != 0;
        const C wt = conj_c(tw[(l & (h - 1)) * q * P]);
        const C w = up ? wt : CT<C>::make(1, 0);
        const F sg = up ? F(-1) : F(1);
#pragma unroll
        for (int p = 0; p < P; p++) {
            const C t = cmul(y[p], w);
            const C o = shfl_xor_c(t, h);
            y[p] = CT<C>::make(o.x + sg * t.x, o.y + sg * t.y);
        }
    }
#pragma unroll
    for (int k1 = 1; k1 < P; k1++) y[k1] = cmul(y[k1], conj_c(tw[l * k1]));
    if constexpr (P > 1) small_dft<P, 1, C>(y);
}

// The same transforms with L fixed at compile time: stages unrolled,
// exchanges by xchg_c (L at most 32).
template <int P, int L, int H, typename C>
__device__ __forceinline__ void fwd_stages(C *y, int l, const C *tw) {
    if constexpr (H >= 1) {
        typedef typename CT<C>::R F;
        const bool up = (l & H) != 0;
        const C wt = tw[(l & (H - 1)) * (L / (2 * H)) * P];
        const C w = up ? wt : CT<C>::make(1, 0);
        const F sg = up ? F(-1) : F(1);
#pragma unroll
        for (int p = 0; p < P; p++) {
            const C o = xchg_c<H>(y[p]);
            y[p] = cmul(CT<C>::make(o.x + sg * y[p].x, o.y + sg * y[p].y), w);
        }
        fwd_stages<P, L, H / 2, C>(y, l, tw);
    }
}
template <int P, int L, int H, typename C>
__device__ __forceinline__ void inv_stages(C *y, int l, const C *tw) {
    if constexpr (H < L) {
        typedef typename CT<C>::R F;
        const bool up = (l & H) != 0;
        const C wt = conj_c(tw[(l & (H - 1)) * (L / (2 * H)) * P]);
        const C w = up ? wt : CT<C>::make(1, 0);
        const F sg = up ? F(-1) : F(1);
#pragma unroll
        for (int p = 0; p < P; p++) {
            const C t = cmul(y[p], w);
            const C o = xchg_c<H>(t);
            y[p] = CT<C>::make(o.x + sg * t.x, o.y + sg * t.y);
        }
        inv_stages<P, L, H * 2, C>(y, l, tw);
    }
}
template <int P, int L, typename C>
__device__ __forceinline__ void fft_fwd_l(C *y, int l, const C *tw) {
    if constexpr (P > 1) small_dft<P, -1, C>(y);
#pragma unroll
    for (int k1 = 1; k1 < P; k1++) y[k1] = cmul(y[k1], tw[l * k1]);
    fwd_stages<P, L, L / 2, C>(y, l, tw);
}
template <int P, int L, typename C>
__device__ __forceinline__ void fft_inv_l(C *y, int l, const C *tw) {
    inv_stages<P, L, 1, C>(y, l, tw);
#pragma unroll
    for (int k1 = 1; k1 < P; k1++) y[k1] = cmul(y[k1], conj_c(tw[l * k1]));
    if constexpr (P > 1) small_dft<P, 1, C>(y);
}
template <int L, typename T> __device__ __forceinline__ T lane_sum(T v) {
    if constexpr (L > 1) {
        v += xchg<L / 2>(v);
        return lane_sum<L / 2>(v);
    } else {
        return v;
    }
}

// Copy of the n twiddles into LDS (before the caller's first barrier).
template <typename C>
__device__ __forceinline__ void load_tw(C *dst, const C *src, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

template <typename C> __device__ __forceinline__ const C *tw_of(CG &G, bool y);
template <> __device__ __forceinline__ const float2 *tw_of<float2>(CG &G, bool y) { return y ? G.twy : G.twx; }
template <> __device__ __forceinline__ const double2 *tw_of<double2>(CG &G, bool y) { return y ? G.twy64 : G.twx64; }

// ---------------------------------------------------------------------------
// TT projection of the RHS gradients: tt = PTT * g per (system, WFS).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_ttr(KARGS) {
    KJOB;
    __shared__ double sh[8];
    const int w = bx, s = J.s0 + bz;
    const float *grad = PD.g;
    float *tt = PD.tt;
    float *out = tt + ((long)s * G.nwfs + w) * 2;
    if (!G.ttr[w]) {
        if (threadIdx.x == 0) out[0] = out[1] = 0.f;
        return;
    }
    const int nsa = G.nsa[w];
    const float *g = grad + (long)s * G.ngtot + G.goff[w];
    const float *P = G.ptt + 4 * G.saoff[w];
    double a0 = 0, a1 = 0;
    for (int i = threadIdx.x; i < 2 * nsa; i += blockDim.x) {
        a0 += (double)P[2 * i] * g[i];
        a1 += (double)P[2 * i + 1] * g[i];
    }
    a0 = block_sum(a0, sh);
    a1 = block_sum(a1, sh);
    if (threadIdx.x == 0) {
        out[0] = (float)a0;
        out[1] = (float)a1;
    }
}

// TT projection of the LHS gradients (TTFR inside TomoL when the LHS removes
// TT, maos/recon_utils.c:461-465): ttl = PTT * (G HX p) per (system, WFS),
// from the gradients k_wfsg wrote, kept in FP64.
template <typename RT>
__global__ __launch_bounds__(256) void k_ttl(KARGS) {
    KJOB;
    __shared__ double sh[8];
    const int w = bx, s = J.s0 + bz;
    double *out = PD.ttl + ((long)s * G.nwfs + w) * 2;
    if (!G.ttr[w]) {
        if (threadIdx.x == 0) out[0] = out[1] = 0.;
        return;
    }
    const int nsa = G.nsa[w];
    const RT *g = (const RT *)PD.gl + (long)s * G.ngtot + G.goff[w];
    const float *P = G.ptt + 4 * G.saoff[w];
    double a0 = 0, a1 = 0;
    for (int i = threadIdx.x; i < 2 * nsa; i += blockDim.x) {
        a0 += (double)P[2 * i] * (double)g[i];
        a1 += (double)P[2 * i + 1] * (double)g[i];
    }
    a0 = block_sum(a0, sh);
    a1 = block_sum(a1, sh);
    if (threadIdx.x == 0) {
        out[0] = a0;
        out[1] = a1;
    }
}

// Gradients of the LHS when it removes TT: gl = GP HX p for the subapertures
// whose lower-left corner lies in this workgroup's TILE x TILE tile of one
// WFS (each subaperture is written by exactly one workgroup).  The TT
// projection (k_ttl) and W GP' (k_wfs<RHS, GL>) follow.
template <int POS, typename RT>
__global__ __launch_bounds__(256) void k_wfsg(KARGS, int xsel) {
    KJOB;
    const float *x = vsel(PD, xsel);
    RT *gl = (RT *)PD.gl;
    constexpr int H = TILE + 2 * POS;
    constexpr int NP = POS + 1;
    __shared__ RT opd[H * H];
    const int w = by, s = J.s0 + bz;
    const int pnx = G.pnx, pny = G.pny;
    const int ntx = (pnx + TILE - 1) / TILE;
    const int tx0 = (bx % ntx) * TILE, ty0 = (bx / ntx) * TILE;
    const float *xs = x + (long)s * G.nxtot;
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) {
        const int px = tx0 - POS + i % H, py = ty0 - POS + i / H;
        RT v = 0;
        if (px >= 0 && px < pnx && py >= 0 && py < pny) {
            for (int l = 0; l < G.npsr; l++) {
                const int wl = w * MAXL + l;
                const int2 nx_ = G.hx_nx[(long)wl * pnx + px];
                const int2 ny_ = G.hx_ny[(long)wl * pny + py];
                if (nx_.x >= 0 && ny_.x >= 0) {
                    const RT fx = G.hx_fx[(long)wl * pnx + px];
                    const RT fy = G.hx_fy[(long)wl * pny + py];
                    const float *xl = xs + G.xoff[l];
                    const int nxl = G.xnx[l];
                    const RT a = xl[ny_.x * nxl + nx_.x], b = xl[ny_.x * nxl + nx_.y];
                    const RT c = xl[ny_.y * nxl + nx_.x], d = xl[ny_.y * nxl + nx_.y];
                    v += (RT(1) - fy) * ((RT(1) - fx) * a + fx * b) + fy * ((RT(1) - fx) * c + fx * d);
                }
            }
        }
        opd[i] = v;
    }
    __syncthreads();
    const int ox = threadIdx.x % TILE, oy = threadIdx.x / TILE;
    const int cx = tx0 + ox, cy = ty0 + oy;
    if (threadIdx.x < TILE * TILE && cx < pnx && cy < pny) {
        const int sa = G.saidx[(long)w * G.nploc + cy * pnx + cx];
        if (sa >= 0) {
            const long so = G.saoff[w] + sa;
            const float *cxp = G.gpx + so * NP * NP, *cyp = G.gpy + so * NP * NP;
            RT gx = 0, gy = 0;
#pragma unroll
            for (int zy = 0; zy < NP; zy++)
#pragma unroll
                for (int zx = 0; zx < NP; zx++) {
                    const RT o = opd[(oy + POS + zy) * H + ox + POS + zx];
                    gx += (RT)cxp[zy * NP + zx] * o;
                    gy += (RT)cyp[zy * NP + zx] * o;
                }
            RT *gw = gl + (long)s * G.ngtot + G.goff[w];
            gw[sa] = gx;
            gw[G.nsa[w] + sa] = gy;
        }
    }
}

// ---------------------------------------------------------------------------
// WFS-side pass.  One workgroup = one TILE x TILE tile of ploc points of one
// WFS.  LHS: opd = HX p on the tile + halo (LDS), g = GP opd, W g, then
// q = GP' (W g) gathered from the <= (pos+1)^2 subapertures touching each
// point.  RHS: W (g - TT tt) then the same GP' gather; with GL the
// gradients are the LHS's own (gl, ttl from k_wfsg and k_ttl) instead of the
// measured ones (g, tt).
// ---------------------------------------------------------------------------
template <int POS, bool RHS, typename RT, bool GL = false>
__global__ __launch_bounds__(256) void k_wfs(KARGS, int xsel) {
    KJOB;
    const float *x = vsel(PD, xsel), *grad = PD.g, *tt = PD.tt;
    const RT *gradl = (const RT *)PD.gl;
    RT *q = (RT *)PD.q;
    constexpr int H = TILE + 2 * POS;
    constexpr int S = TILE + POS;
    constexpr int NP = POS + 1;
    __shared__ RT opd[RHS ? 1 : H * H];
    __shared__ RT wgx[S * S], wgy[S * S];
    __shared__ int sidx[S * S];
    const int w = by, s = J.s0 + bz;
    const int pnx = G.pnx, pny = G.pny;
    const int ntx = (pnx + TILE - 1) / TILE;
    const int tx0 = (bx % ntx) * TILE, ty0 = (bx / ntx) * TILE;
    if (!RHS) {
        const float *xs = x + (long)s * G.nxtot;
        for (int i = threadIdx.x; i < H * H; i += blockDim.x) {
            const int px = tx0 - POS + i % H, py = ty0 - POS + i / H;
            RT v = 0;
            if (px >= 0 && px < pnx && py >= 0 && py < pny) {
                for (int l = 0; l < G.npsr; l++) {
                    const int wl = w * MAXL + l;
                    const int2 nx_ = G.hx_nx[(long)wl * pnx + px];
                    const int2 ny_ = G.hx_ny[(long)wl * pny + py];
                    if (nx_.x >= 0 && ny_.x >= 0) {
                        const RT fx = G.hx_fx[(long)wl * pnx + px];
                        const RT fy = G.hx_fy[(long)wl * pny + py];
                        const float *xl = xs + G.xoff[l];
                        const int nxl = G.xnx[l];
                        const RT a = xl[ny_.x * nxl + nx_.x], b = xl[ny_.x * nxl + nx_.y];
                        const RT c = xl[ny_.y * nxl + nx_.x], d = xl[ny_.y * nxl + nx_.y];
                        v += (RT(1) - fy) * ((RT(1) - fx) * a + fx * b) + fy * ((RT(1) - fx) * c + fx * d);
                    }
                }
            }
            opd[i] = v;
        }
        __syncthreads();
    }
    const int nsa = G.nsa[w];
    const float *gw = grad + (long)s * G.ngtot + G.goff[w];
    const RT *gwl = GL ? gradl + (long)s * G.ngtot + G.goff[w] : nullptr;
    RT ttx = 0, tty = 0;
    if (RHS && G.ttr[w]) {
        if (GL) {
            ttx = (RT)PD.ttl[((long)s * G.nwfs + w) * 2];
            tty = (RT)PD.ttl[((long)s * G.nwfs + w) * 2 + 1];
        } else {
            ttx = tt[((long)s * G.nwfs + w) * 2];
            tty = tt[((long)s * G.nwfs + w) * 2 + 1];
        }
    }
    for (int i = threadIdx.x; i < S * S; i += blockDim.x) {
        const int a = i % S, b = i / S;
        const int cx = tx0 - POS + a, cy = ty0 - POS + b;
        int sa = -1;
        if (cx >= 0 && cx < pnx && cy >= 0 && cy < pny) sa = G.saidx[(long)w * G.nploc + cy * pnx + cx];
        sidx[i] = sa;
        RT ox = 0, oy = 0;
        if (sa >= 0) {
            const long so = G.saoff[w] + sa;
            RT gx = 0, gy = 0;
            if (RHS && GL) {
                gx = gwl[sa] - ttx;
                gy = gwl[nsa + sa] - tty;
            } else if (RHS) {
                gx = (RT)gw[sa] - ttx;
                gy = (RT)gw[nsa + sa] - tty;
            } else {
                const float *cxp = G.gpx + so * NP * NP, *cyp = G.gpy + so * NP * NP;
#pragma unroll
                for (int zy = 0; zy < NP; zy++)
#pragma unroll
                    for (int zx = 0; zx < NP; zx++) {
                        const RT o = opd[(b + zy) * H + a + zx];
                        gx += (RT)cxp[zy * NP + zx] * o;
                        gy += (RT)cyp[zy * NP + zx] * o;
                    }
            }
            const float4 ne = G.neai[so];
            ox = (RT)ne.x * gx + (RT)ne.z * gy;
            oy = (RT)ne.z * gx + (RT)ne.y * gy;
        }
        wgx[i] = ox;
        wgy[i] = oy;
    }
    __syncthreads();
    const int ox = threadIdx.x % TILE, oy = threadIdx.x / TILE;
    const int px = tx0 + ox, py = ty0 + oy;
    if (threadIdx.x < TILE * TILE && px < pnx && py < pny) {
        RT acc = 0;
#pragma unroll
        for (int zy = 0; zy < NP; zy++)
#pragma unroll
            for (int zx = 0; zx < NP; zx++) {
                const int si = (oy + POS - zy) * S + (ox + POS - zx);
                const int sa = sidx[si];
                if (sa >= 0) {
                    const long so = G.saoff[w] + sa;
                    acc += (RT)G.gpx[so * NP * NP + zy * NP + zx] * wgx[si] +
                           (RT)G.gpy[so * NP * NP + zy * NP + zx] * wgy[si];
                }
            }
        q[((long)s * G.nwfs + w) * G.nploc + py * pnx + px] = acc;
    }
}

// ---------------------------------------------------------------------------
// Layer-side pass.  out = sum_w HX_wl' q_w  (+ L2'L2 p for the LHS), gathered
// per layer node from the <= TAPS^2 ploc samples whose bilinear footprint
// covers it.  Also writes the FP64 partial of p.Ap (LHS) or out.out (RHS).
// ---------------------------------------------------------------------------
template <bool LHS, typename RT>
__global__ __launch_bounds__(256) void k_layer(KARGS, int psel_, int osel, int partsel) {
    KJOB;
    const RT *q = (const RT *)PD.q;
    const float *p = vsel(PD, psel_);
    RT *out = rsel<RT>(PD, osel);
    double *part = psel(PD, partsel);
    const int partstride = PD.SC.maxpart;
    __shared__ double sh[8];
    const int l = by, s = J.s0 + bz;
    const int nxl = G.xnx[l], nyl = G.xny[l];
    const int ntx = (nxl + TILE - 1) / TILE, nty = (nyl + TILE - 1) / TILE;
    double dot = 0;
    if (bx < ntx * nty) {
        const int jx = (bx % ntx) * TILE + threadIdx.x % TILE;
        const int jy = (bx / ntx) * TILE + threadIdx.x / TILE;
        if (jx < nxl && jy < nyl) {
            RT acc = 0;
            const long xt = G.hxt_xoff[G.npsr];  // per-WFS stride of transposed tables
            for (int w = 0; w < G.nwfs; w++) {
                const int wl = w * MAXL + l;
                const int4 cx = G.hxt_x[(long)w * xt + G.hxt_xoff[l] + jx];
                const int4 cy = G.hxt_y[(long)w * xt + G.hxt_xoff[l] + jy];
                const int ix[TAPS] = {cx.x, cx.y, cx.z, cx.w};
                const int iy[TAPS] = {cy.x, cy.y, cy.z, cy.w};
                float wx[TAPS], wy[TAPS];
#pragma unroll
                for (int k = 0; k < TAPS; k++) {
                    wx[k] = 0.f;
                    wy[k] = 0.f;
                    if (ix[k] >= 0) {
                        const int2 nn = G.hx_nx[(long)wl * G.pnx + ix[k]];
                        const float f = G.hx_fx[(long)wl * G.pnx + ix[k]];
                        wx[k] = (nn.x == jx ? 1.f - f : 0.f) + (nn.y == jx ? f : 0.f);
                    }
                    if (iy[k] >= 0) {
                        const int2 nn = G.hx_ny[(long)wl * G.pny + iy[k]];
                        const float f = G.hx_fy[(long)wl * G.pny + iy[k]];
                        wy[k] = (nn.x == jy ? 1.f - f : 0.f) + (nn.y == jy ? f : 0.f);
                    }
                }
                const RT *qw = q + ((long)s * G.nwfs + w) * G.nploc;
#pragma unroll
                for (int ky = 0; ky < TAPS; ky++) {
                    if (iy[ky] < 0) continue;
                    RT ra = 0;
#pragma unroll
                    for (int kx = 0; kx < TAPS; kx++)
                        if (ix[kx] >= 0) ra += (RT)wx[kx] * qw[iy[ky] * G.pnx + ix[kx]];
                    acc += (RT)wy[ky] * ra;
                }
            }
            const long o = (long)s * G.nxtot + G.xoff[l];
            if (LHS) {
                // periodic L2'L2 stencil (20, -8, 2, 1) * l2c, accumulated in RT:
                // its terms are large and cancel for smooth p
                const float *pl = p + o;
                auto P = [&](int dx, int dy) {
                    int xx = jx + dx, yy = jy + dy;
                    xx = xx < 0 ? xx + nxl : (xx >= nxl ? xx - nxl : xx);
                    yy = yy < 0 ? yy + nyl : (yy >= nyl ? yy - nyl : yy);
                    return (RT)pl[yy * nxl + xx];
                };
                const RT c = P(0, 0);
                const RT s4 = P(-1, 0) + P(1, 0) + P(0, -1) + P(0, 1);
                const RT sd = P(-1, -1) + P(1, -1) + P(-1, 1) + P(1, 1);
                const RT s2 = P(-2, 0) + P(2, 0) + P(0, -2) + P(0, 2);
                acc += (RT)G.l2c[l] * (RT(20) * c - RT(8) * s4 + RT(2) * sd + s2);
                dot = dmul(c, acc, G.dot32);
            } else {
                dot = dmul(acc, acc, G.dot32);
            }
            out[o + jy * nxl + jx] = acc;
        }
    }
    dot = block_sum(dot, sh, G.dot32);
    if (threadIdx.x == 0 && part) part[(long)s * partstride + by * PD.ltiles + bx] = dot;
}

// ---------------------------------------------------------------------------
// Matched-sampling fast path (ploc spacing * cone scale == layer spacing for
// every WFS and layer: LGS in cone coordinates, NGS on an NGS-only system).
// The ray of (w, l) is then a constant integer offset plus a constant
// bilinear fraction, so HX is a 2x2 stencil on a shifted window and HX' its
// mirror.  Layer windows (k_wfs_m) and WFS-field windows (k_layer_m) are
// staged in LDS with coalesced loads.
// ---------------------------------------------------------------------------
template <int POS, typename RT, int CS>
__global__ __launch_bounds__(256) void k_wfs_m(KARGS, int xsel) {
    KJOB;
    const float *x = vsel(PD, xsel);
    RT *q = (RT *)PD.q;
    constexpr int H = TILE + 2 * POS;   // opd window edge
    constexpr int S = TILE + POS;       // subaperture-corner window edge
    constexpr int S2 = S * S;
    constexpr int NP = POS + 1, NP2 = NP * NP;
    constexpr int UG = (S2 * NP2 + 255) / 256, UN = (S2 + 255) / 256;
    static_assert(H % 2 == 0 && H * H / 2 <= 256, "one two-row item per thread");
    __shared__ float2 gps[S2 * NP2];  // GP coefficients of the window's corners
    __shared__ RT opd[CS][H * H];
    __shared__ RT wgx[CS][S2], wgy[CS][S2];
    const int w = by, s0 = J.s0 + bz * CS, ns = min(CS, J.s0 + J.nsys - s0);
    const int pnx = G.pnx, pny = G.pny;
    const int ntx = (pnx + TILE - 1) / TILE;
    const int tx0 = (bx % ntx) * TILE, ty0 = (bx / ntx) * TILE;
    const float2 *gpw = G.gpd + (long)G.gcls[w] * G.nploc * NP2;
    const float4 *nw4 = G.nead + (long)G.gcls[w] * G.nploc;
    // geometry of the window (shared by the CS systems).  Loads here and in
    // the HX pass are unconditional, from indices clamped into range, and
    // masked afterwards: loads under a branch are each waited for before
    // the branch closes, which would serialise them.
    float2 gv[UG];
    float4 nv[UN];
#pragma unroll
    for (int u = 0; u < UG; u++) {
        const int i = min((int)threadIdx.x + 256 * u, S2 * NP2 - 1);
        const int c = i / NP2, z = i - c * NP2;
        const int cx = tx0 - POS + c % S, cy = ty0 - POS + c / S;
        const int cxc = min(max(cx, 0), pnx - 1), cyc = min(max(cy, 0), pny - 1);
        const float2 v = gpw[(long)(cyc * pnx + cxc) * NP2 + z];
        gv[u] = (cx == cxc && cy == cyc) ? v : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < UN; u++) {
        const int i = threadIdx.x + 256 * u;
        const int ic = min(i, S2 - 1);
        const int cx = tx0 - POS + ic % S, cy = ty0 - POS + ic / S;
        const int cxc = min(max(cx, 0), pnx - 1), cyc = min(max(cy, 0), pny - 1);
        const float4 v = nw4[cyc * pnx + cxc];
        nv[u] = (i < S2 && cx == cxc && cy == cyc) ? v : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < UG; u++) {
        // past the end: the last entry again, with its own value
        const int i = min((int)threadIdx.x + 256 * u, S2 * NP2 - 1);
        gps[i] = gv[u];
    }
    // HX on the opd window, gathered straight from the layers (L1/L2 hits):
    // ploc point i reads nodes i+c and i+c+1 (the latter wrapping to 0 at the
    // last node, the reference's edge ray).  Per layer the window is either
    // wholly inside the ray interval (plain loads) or not (indices clamped
    // into the grid, points outside the interval contribute zero).  Index
    // work is shared by the CS systems.
    // Work item i: column i % H of the window, rows 2*(i / H) and +1 (H is
    // even), so on the plain path the middle sample row serves both points.
    constexpr int NI = H * H / 2;
    const int ia = threadIdx.x % H, ib = threadIdx.x / H;
    const bool item = threadIdx.x < NI;
    const int hx = tx0 - POS + ia, hy0 = ty0 - POS + 2 * ib;
    RT v[2][CS];
#pragma unroll
    for (int r = 0; r < 2; r++)
#pragma unroll
        for (int si = 0; si < CS; si++) v[r][si] = 0;
    if (item) {
        const float *xs = x + (long)s0 * G.nxtot;
        for (int l = 0; l < G.npsr; l++) {
            const MRay m = mray(G, w * MAXL + l);
            const int nxl = G.xnx[l], nyl = G.xny[l];
            const long xo = G.xoff[l];
            // every scalar of this layer in one round trip (the compiler would
            // otherwise load those used only past a branch separately)
            asm volatile("" ::"s"(m.wx), "s"(m.wy), "s"(nyl), "s"(xo));
            const float *xl = xs + xo;
            const RT fx = m.fx, fy = m.fy;
            const bool fast = tx0 - POS >= m.ax0 && tx0 + TILE + POS <= (m.wx ? m.ax1 - 1 : m.ax1) &&
                              ty0 - POS >= m.ay0 && ty0 + TILE + POS <= (m.wy ? m.ay1 - 1 : m.ay1);
            // systems past ns read system ns-1 (in range; never stored)
            if (fast) {
                const int o = (hy0 + m.cy) * nxl + (hx + m.cx);
#pragma unroll
                for (int si = 0; si < CS; si++) {
                    const float *t = xl + (long)min(si, ns - 1) * G.nxtot + o;
                    RT h[3];
#pragma unroll
                    for (int k = 0; k < 3; k++) h[k] = (RT(1) - fx) * (RT)t[k * nxl] + fx * (RT)t[k * nxl + 1];
                    v[0][si] += (RT(1) - fy) * h[0] + fy * h[1];
                    v[1][si] += (RT(1) - fy) * h[1] + fy * h[2];
                }
            } else {
#pragma unroll
                for (int r = 0; r < 2; r++) {
                    const int hy = hy0 + r;
                    const bool ok = hx >= m.ax0 && hx < m.ax1 && hy >= m.ay0 && hy < m.ay1;
                    const int n0x = min(max(hx + m.cx, 0), nxl - 1), n0y = min(max(hy + m.cy, 0), nyl - 1);
                    const int n1x = n0x + 1 == nxl ? 0 : n0x + 1;
                    const int n1y = n0y + 1 == nyl ? 0 : n0y + 1;
                    const int o00 = n0y * nxl + n0x, o10 = n0y * nxl + n1x, o01 = n1y * nxl + n0x,
                              o11 = n1y * nxl + n1x;
#pragma unroll
                    for (int si = 0; si < CS; si++) {
                        const float *t = xl + (long)min(si, ns - 1) * G.nxtot;
                        const RT a = t[o00], b = t[o10], c = t[o01], d = t[o11];
                        const RT h = (RT(1) - fy) * ((RT(1) - fx) * a + fx * b) + fy * ((RT(1) - fx) * c + fx * d);
                        v[r][si] += ok ? h : RT(0);
                    }
                }
            }
        }
#pragma unroll
        for (int r = 0; r < 2; r++)
#pragma unroll
            for (int si = 0; si < CS; si++)
                if (si < ns) opd[si][(2 * ib + r) * H + ia] = v[r][si];
    }
    __syncthreads();
    // W GP per subaperture corner of the window (zero coefficients where none)
#pragma unroll
    for (int u = 0; u < UN; u++) {
        const int i = threadIdx.x + 256 * u;
        if (i < S2) {
            const int a = i % S, b = i / S;
            float2 c[NP2];
#pragma unroll
            for (int z = 0; z < NP2; z++) c[z] = gps[i * NP2 + z];
            const float4 ne = nv[u];
#pragma unroll
            for (int si = 0; si < CS; si++) {
                if (si < ns) {
                    RT gx = 0, gy = 0;
#pragma unroll
                    for (int zy = 0; zy < NP; zy++)
#pragma unroll
                        for (int zx = 0; zx < NP; zx++) {
                            const RT o = opd[si][(b + zy) * H + a + zx];
                            gx += (RT)c[zy * NP + zx].x * o;
                            gy += (RT)c[zy * NP + zx].y * o;
                        }
                    wgx[si][i] = (RT)ne.x * gx + (RT)ne.z * gy;
                    wgy[si][i] = (RT)ne.z * gx + (RT)ne.y * gy;
                }
            }
        }
    }
    __syncthreads();
    // GP' gather onto the tile
    const int ox = threadIdx.x % TILE, oy = threadIdx.x / TILE;
    const int px = tx0 + ox, py = ty0 + oy;
    if (threadIdx.x < TILE * TILE && px < pnx && py < pny) {
        float2 c[NP2];
        int sidx[NP2];
#pragma unroll
        for (int zy = 0; zy < NP; zy++)
#pragma unroll
            for (int zx = 0; zx < NP; zx++) {
                const int si = (oy + POS - zy) * S + (ox + POS - zx);
                sidx[zy * NP + zx] = si;
                c[zy * NP + zx] = gps[si * NP2 + zy * NP + zx];
            }
#pragma unroll
        for (int si = 0; si < CS; si++) {
            if (si < ns) {
                RT acc = 0;
#pragma unroll
                for (int z = 0; z < NP2; z++) acc += (RT)c[z].x * wgx[si][sidx[z]] + (RT)c[z].y * wgy[si][sidx[z]];
                q[((long)(s0 + si) * G.nwfs + w) * G.nploc + py * pnx + px] = acc;
            }
        }
    }
}

template <bool LHS, typename RT, int CS>
__global__ __launch_bounds__(256) void k_layer_m(KARGS, int psel_, int osel, int partsel) {
    KJOB;
    const RT *q = (const RT *)PD.q;
    const float *p = vsel(PD, psel_);
    RT *out = rsel<RT>(PD, osel);
    double *part = psel(PD, partsel);
    const int partstride = PD.SC.maxpart;
    constexpr int PH = TILE + 4;     // p window edge (L2'L2 reach is 2)
    constexpr int PH2 = PH * PH;
    constexpr int UP = (PH2 + 255) / 256;
    __shared__ float ps[CS][LHS ? PH2 : 1];
    __shared__ double sh[8];
    const int l = by, s0 = J.s0 + bz * CS, ns = min(CS, J.s0 + J.nsys - s0);
    const int nxl = G.xnx[l], nyl = G.xny[l];
    const int ntx = (nxl + TILE - 1) / TILE, nty = (nyl + TILE - 1) / TILE;
    const bool active = bx < ntx * nty;
    const int jx0 = (bx % ntx) * TILE, jy0 = (bx / ntx) * TILE;
    const int jx = jx0 + (int)threadIdx.x % TILE, jy = jy0 + (int)threadIdx.x / TILE;
    const bool mine = active && jx < nxl && jy < nyl;
    if (LHS && active) {
        // p windows, periodic (layers are >= TILE + 2, so one wrap suffices)
        int off[UP];
#pragma unroll
        for (int u = 0; u < UP; u++) {
            const int i = threadIdx.x + 256 * u;
            int xx = jx0 - 2 + i % PH, yy = jy0 - 2 + i / PH;
            xx += xx < 0 ? nxl : 0;
            xx -= xx >= nxl ? nxl : 0;
            yy += yy < 0 ? nyl : 0;
            yy -= yy >= nyl ? nyl : 0;
            off[u] = yy * nxl + xx;
        }
        float pv[CS][UP];
#pragma unroll
        for (int si = 0; si < CS; si++) {
            const float *pl = p + (long)(s0 + si) * G.nxtot + G.xoff[l];
#pragma unroll
            for (int u = 0; u < UP; u++) {
                const int i = threadIdx.x + 256 * u;
                pv[si][u] = (si < ns && i < PH2) ? pl[off[u]] : 0.f;
            }
        }
#pragma unroll
        for (int si = 0; si < CS; si++)
#pragma unroll
            for (int u = 0; u < UP; u++) {
                const int i = threadIdx.x + 256 * u;
                if (i < PH2) ps[si][i] = pv[si][u];
            }
    }
    // HX' gather straight from the WFS fields (L1/L2 hits): node j takes the
    // ray whose second node it is (ploc index j-c-1, weight f) and the ray
    // whose first node it is (j-c, weight 1-f) on each axis.  Per WFS the tile
    // is classified uniformly: no ray lands on it (skip), every node's four
    // rays are valid (plain loads), or mixed (indices clamped into the grid,
    // rays outside the interval weighted zero).  Index work is shared by the
    // CS systems.
    RT acc[CS];
#pragma unroll
    for (int si = 0; si < CS; si++) acc[si] = 0;
    const int tw = min(TILE, nxl - jx0), th = min(TILE, nyl - jy0);
    if (mine) {
        for (int w = 0; w < G.nwfs; w++) {
            const MRay m = mray(G, w * MAXL + l);
            const bool edgex = m.wx && jx0 == 0, edgey = m.wy && jy0 == 0;
            if (!edgex && (jx0 + tw - 1 - m.cx < m.ax0 || jx0 - m.cx - 1 >= m.ax1)) continue;
            if (!edgey && (jy0 + th - 1 - m.cy < m.ay0 || jy0 - m.cy - 1 >= m.ay1)) continue;
            const RT *qw = q + ((long)s0 * G.nwfs + w) * G.nploc;
            const long qstride = (long)G.nwfs * G.nploc;
            const RT fx = m.fx, fy = m.fy;
            if (!edgex && !edgey && jx0 - m.cx - 1 >= m.ax0 && jx0 + tw - m.cx <= m.ax1 &&
                jy0 - m.cy - 1 >= m.ay0 && jy0 + th - m.cy <= m.ay1) {
                const int o = (jy - m.cy - 1) * G.pnx + (jx - m.cx - 1);
#pragma unroll
                for (int si = 0; si < CS; si++) {
                    if (si < ns) {
                        const RT *t = qw + si * qstride + o;
                        const RT q00 = t[0], q10 = t[1], q01 = t[G.pnx], q11 = t[G.pnx + 1];
                        acc[si] += fy * ((RT(1) - fx) * q10 + fx * q00) + (RT(1) - fy) * ((RT(1) - fx) * q11 + fx * q01);
                    }
                }
                continue;
            }
            int ix0 = jx - m.cx - 1, iy0 = jy - m.cy - 1;
            // node 0 is the wrapped second node of the edge ray a1 - 1
            if (m.wx && jx == 0) ix0 = m.ax1 - 1;
            if (m.wy && jy == 0) iy0 = m.ay1 - 1;
            const int ix1 = jx - m.cx, iy1 = jy - m.cy;
            const RT wx0 = (ix0 >= m.ax0 && ix0 < m.ax1) ? fx : RT(0);
            const RT wx1 = (ix1 >= m.ax0 && ix1 < m.ax1) ? RT(1) - fx : RT(0);
            const RT wy0 = (iy0 >= m.ay0 && iy0 < m.ay1) ? fy : RT(0);
            const RT wy1 = (iy1 >= m.ay0 && iy1 < m.ay1) ? RT(1) - fy : RT(0);
            const int cx0 = min(max(ix0, 0), G.pnx - 1), cx1 = min(max(ix1, 0), G.pnx - 1);
            const int cy0 = min(max(iy0, 0), G.pny - 1), cy1 = min(max(iy1, 0), G.pny - 1);
            const int o00 = cy0 * G.pnx + cx0, o10 = cy0 * G.pnx + cx1, o01 = cy1 * G.pnx + cx0,
                      o11 = cy1 * G.pnx + cx1;
#pragma unroll
            for (int si = 0; si < CS; si++) {
                if (si < ns) {
                    const RT *t = qw + si * qstride;
                    acc[si] += wy0 * (wx1 * t[o10] + wx0 * t[o00]) + wy1 * (wx1 * t[o11] + wx0 * t[o01]);
                }
            }
        }
    }
    if (LHS) __syncthreads();
    double dot[CS];
#pragma unroll
    for (int si = 0; si < CS; si++) {
        dot[si] = 0;
        if (mine && si < ns) {
            RT a = acc[si];
            if (LHS) {
                const int tx = threadIdx.x % TILE, ty = threadIdx.x / TILE;
                const float *P0 = ps[si] + (ty + 2) * PH + tx + 2;
                auto P = [&](int dx, int dy) { return (RT)P0[dy * PH + dx]; };
                const RT c = P(0, 0);
                const RT s4 = P(-1, 0) + P(1, 0) + P(0, -1) + P(0, 1);
                const RT sd = P(-1, -1) + P(1, -1) + P(-1, 1) + P(1, 1);
                const RT s2 = P(-2, 0) + P(2, 0) + P(0, -2) + P(0, 2);
                a += (RT)G.l2c[l] * (RT(20) * c - RT(8) * s4 + RT(2) * sd + s2);
                dot[si] = dmul(c, a, G.dot32);
            } else {
                dot[si] = dmul(a, a, G.dot32);
            }
            out[(long)(s0 + si) * G.nxtot + G.xoff[l] + jy * nxl + jx] = a;
        }
    }
#pragma unroll
    for (int si = 0; si < CS; si++) {
        const double d = block_sum(dot[si], sh, G.dot32);
        if (threadIdx.x == 0 && part && si < ns)
            part[(long)(s0 + si) * partstride + by * PD.ltiles + bx] = d;
    }
}

// Layer pass of the matched path on 16 x LT tiles (LT = 64 or 32) with 128
// threads: each thread owns two adjacent columns of R = LT / 16 vertically
// adjacent nodes, so the R + 1 sample rows of a WFS field, read at three ray
// columns (the middle one serves both columns), feed 2 R nodes.  Same arithmetic per node as
// k_layer_m (the reference's bilinear transpose):
//   acc += fy (gx q[iy0, ix1] + fx q[iy0, ix0]) + gy (gx q[iy1, ix1] + fx q[iy1, ix0]),
// gx = 1 - fx, gy = 1 - fy; the p.Ap partial of a thread sums its nodes.
template <bool LHS, typename RT, int CS, int LT>
__global__ __launch_bounds__(128) void k_layer_m8(KARGS, int psel_, int osel, int partsel) {
    KJOB;
    const RT *q = (const RT *)PD.q;
    const float *p = vsel(PD, psel_);
    RT *out = rsel<RT>(PD, osel);
    double *part = psel(PD, partsel);
    const int partstride = PD.SC.maxpart;
    constexpr int R = LT / 16, NT = 128;
    constexpr int PW = TILE + 4, PHH = LT + 4, P2 = PW * PHH;  // p window (L2'L2 reach 2)
    constexpr int UP = (P2 + NT - 1) / NT;
    __shared__ float ps[CS][LHS ? P2 : 1];
    __shared__ double sh[8];
    const int l = by, s0 = J.s0 + bz * CS, ns = min(CS, J.s0 + J.nsys - s0);
    const int nxl = G.xnx[l], nyl = G.xny[l];
    const int ntx = (nxl + TILE - 1) / TILE, nty = (nyl + LT - 1) / LT;
    const bool active = bx < ntx * nty;
    const int jx0 = (bx % ntx) * TILE, jy0 = (bx / ntx) * LT;
    const int tx = threadIdx.x % (TILE / 2), ty = threadIdx.x / (TILE / 2);
    const int jx = jx0 + 2 * tx, jyb = jy0 + R * ty;
    const int nr = active ? max(0, min(R, nyl - jyb)) : 0;  // rows of this thread
    const int nc = active ? max(0, min(2, nxl - jx)) : 0;   // columns of this thread
    if (LHS && active) {
        // p windows, periodic (layers are >= LT + 2 rows and TILE + 2 columns)
        int off[UP];
#pragma unroll
        for (int u = 0; u < UP; u++) {
            const int i = min((int)threadIdx.x + NT * u, P2 - 1);
            int xx = jx0 - 2 + i % PW, yy = jy0 - 2 + i / PW;
            xx += xx < 0 ? nxl : 0;
            xx -= xx >= nxl ? nxl : 0;
            yy += yy < 0 ? nyl : 0;
            yy -= yy >= nyl ? nyl : 0;
            off[u] = yy * nxl + xx;
        }
        float pv[CS][UP];
#pragma unroll
        for (int si = 0; si < CS; si++) {
            const float *pl = p + (long)(s0 + min(si, ns - 1)) * G.nxtot + G.xoff[l];
#pragma unroll
            for (int u = 0; u < UP; u++) pv[si][u] = pl[off[u]];
        }
#pragma unroll
        for (int si = 0; si < CS; si++)
#pragma unroll
            for (int u = 0; u < UP; u++) ps[si][min((int)threadIdx.x + NT * u, P2 - 1)] = pv[si][u];
    }
    RT acc[CS][2][R];
#pragma unroll
    for (int si = 0; si < CS; si++)
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int r = 0; r < R; r++) acc[si][c][r] = 0;
    const int tw = min(TILE, nxl - jx0), th = min(LT, nyl - jy0);
    // Node jx takes rays ix0 = jx - cx - 1 (the last ray for node 0 on a
    // wrapped axis) and ix1 = jx - cx; node jx + 1 takes ix1 and ix1 + 1.
    // Sample rows k = 0..R: iy = jyb - cy - 1 + k (row 0: the last ray when
    // jyb = 0 wraps).  The buffer of this WFS starts at the first valid row
    // ay0 and holds the valid rows [ay0, ay1) only, so a sample row outside
    // them falls outside it (rows above wrap to offsets past its end) and
    // reads 0 with no test per row; an invalid ray column starts from BOOB,
    // past the end for every row.
    if (nr > 0 && nc > 0) {
        const unsigned rstep = (unsigned)G.pnx * (unsigned)sizeof(RT);
        const RT *qs[CS];  // the WFS fields of each system
#pragma unroll
        for (int si = 0; si < CS; si++) qs[si] = q + (long)(s0 + min(si, ns - 1)) * G.nwfs * G.nploc;
        for (int w = 0; w < G.nwfs; w++) {
            const MRay m = mray(G, w * MAXL + l);
            // the whole record in one scalar round trip (the compiler would
            // otherwise load the fields used after the tile test separately)
            asm volatile("" ::"s"(m.fx), "s"(m.fy), "s"(m.nrow), "s"(m.qoff));
            const bool edgex = m.wx && jx0 == 0, edgey = m.wy && jy0 == 0;
            if (!edgex && (jx0 + tw - 1 - m.cx < m.ax0 || jx0 - m.cx - 1 >= m.ax1)) continue;
            if (!edgey && (jy0 + th - 1 - m.cy < m.ay0 || jy0 - m.cy - 1 >= m.ay1)) continue;
            const RT fx = m.fx, fy = m.fy, gx = RT(1) - fx, gy = RT(1) - fy;
            const int ixa = (m.wx && jx == 0) ? m.ax1 - 1 : jx - m.cx - 1, ixb = jx - m.cx, ixc = ixb + 1;
            const bool va = ixa >= m.ax0 && ixa < m.ax1, vb = ixb >= m.ax0 && ixb < m.ax1,
                       vc = ixc >= m.ax0 && ixc < m.ax1;
            const unsigned ca = va ? (unsigned)ixa * (unsigned)sizeof(RT) : BOOB;
            const unsigned cb = vb ? (unsigned)ixb * (unsigned)sizeof(RT) : BOOB;
            const unsigned cc = vc ? (unsigned)ixc * (unsigned)sizeof(RT) : BOOB;
            const int iyw = (m.wy && jyb == 0) ? m.ay1 - 1 : jyb - m.cy - 1;  // sample row 0
            const unsigned r1 = (unsigned)(jyb - m.cy - m.ay0) * rstep;
            unsigned ro[R + 1];
            ro[0] = (unsigned)(iyw - m.ay0) * rstep;
#pragma unroll
            for (int k = 1; k <= R; k++) ro[k] = r1 + (unsigned)(k - 1) * rstep;
            const unsigned nrec = (unsigned)m.nrow * rstep;
#pragma unroll
            for (int si = 0; si < CS; si++) {
                const auto rs = brsrc(qs[si] + m.qoff, nrec);
                RT h0[R + 1], h1[R + 1];
#pragma unroll
                for (int k = 0; k <= R; k++) {
                    const RT qa = bload<RT>(rs, ca + ro[k]), qb = bload<RT>(rs, cb + ro[k]),
                             qc = bload<RT>(rs, cc + ro[k]);
                    h0[k] = gx * qb + fx * qa;
                    h1[k] = gx * qc + fx * qb;
                }
#pragma unroll
                for (int r = 0; r < R; r++) {
                    acc[si][0][r] += fy * h0[r] + gy * h0[r + 1];
                    acc[si][1][r] += fy * h1[r] + gy * h1[r + 1];
                }
            }
        }
    }
    if (LHS) __syncthreads();
    double dot[CS];
#pragma unroll
    for (int si = 0; si < CS; si++) {
        dot[si] = 0;
        if (si < ns) {
#pragma unroll
            for (int c = 0; c < 2; c++) {
                if (c >= nc) continue;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    if (r >= nr) continue;
                    const int jy = jyb + r;
                    RT a = acc[si][c][r];
                    if (LHS) {
                        const float *P0 = ps[si] + (R * ty + r + 2) * PW + 2 * tx + c + 2;
                        auto P = [&](int dx, int dy) { return (RT)P0[dy * PW + dx]; };
                        const RT cv = P(0, 0);
                        const RT s4 = P(-1, 0) + P(1, 0) + P(0, -1) + P(0, 1);
                        const RT sd = P(-1, -1) + P(1, -1) + P(-1, 1) + P(1, 1);
                        const RT s2 = P(-2, 0) + P(2, 0) + P(0, -2) + P(0, 2);
                        a += (RT)G.l2c[l] * (RT(20) * cv - RT(8) * s4 + RT(2) * sd + s2);
                        dot[si] = dfma(cv, a, dot[si], G.dot32);
                    } else {
                        dot[si] = dfma(a, a, dot[si], G.dot32);
                    }
                    out[(long)(s0 + si) * G.nxtot + G.xoff[l] + jy * nxl + jx + c] = a;
                }
            }
        }
    }
#pragma unroll
    for (int si = 0; si < CS; si++) {
        const double d = block_sum(dot[si], sh, G.dot32);
        if (threadIdx.x == 0 && part && si < ns)
            part[(long)(s0 + si) * partstride + by * PD.ltiles + bx] = d;
    }
}

// ---------------------------------------------------------------------------
// Row kernel: PCG vector update fused with the forward real-to-complex row
// FFT of r.  Two real rows are packed into one complex FFT.  RP = 0: the FFT
// in LDS stages; RP = 2 or 3 (rows of n = 32 RP): in registers, 32 lanes per
// row pair (lane l: columns l + 32 p), written to LDS in bin order.
// ---------------------------------------------------------------------------
enum { ROW_INIT_COLD = 0, ROW_INIT_WARM = 1, ROW_UPDATE = 2, ROW_LAST = 3 };

template <int MODE, typename RT, typename FT, int RM, int RP>
__global__ __launch_bounds__(256) void k_row(KARGS, int iter, int xsel) {
    KJOB;
    const Scal SC = ldsc(PD.SC);
    float *x = vsel(PD, xsel);
    RT *r = (RT *)PD.r;
    const RT *b = (const RT *)PD.b, *Ap = (const RT *)PD.Ap;
    const float *p = PD.p;
    FT *spec = (FT *)PD.specf;
    int *stop = PD.stop;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    FT *lds = reinterpret_cast<FT *>(lds_raw);
    typedef typename CT<FT>::R FR;
    __shared__ double sh[8];
    const int s = J.s0 + bz;
    const int2 rb = G.rowblk[bx];
    const int l = rb.x, row0 = rb.y;
    const int n = G.fnx, nyl = G.xny[l];
    const int nrows = min(ROWS, nyl - row0);
    const int nf = nrows / 2;
    FT *twl = lds + ROWS * n;
    if (MODE != ROW_LAST && RP == 0) load_tw(twl, tw_of<FT>(G, false), n);
    const long base = (long)s * G.nxtot + G.xoff[l] + (long)row0 * n;
    const int tot = nrows * n;
    // register path: lane group f takes rows 2f, 2f + 1 (groups past the
    // last pair repeat it and store nothing), lane lr columns lr + 32 p
    constexpr int RPL = RP > 0 ? RP : 1;
    const int fg = threadIdx.x >> 5, lr = threadIdx.x & 31;
    const int fc = min(fg, nf - 1);
    RT w0[2][RPL], w1[2][RPL];
    if constexpr (RP > 0) {
#pragma unroll
        for (int h = 0; h < 2; h++)
#pragma unroll
            for (int q = 0; q < RP; q++) {
                const long idx = base + (long)(2 * fc + h) * n + lr + 32 * q;
                if (MODE == ROW_INIT_COLD) {
                    w0[h][q] = b[idx];
                } else {
                    w0[h][q] = MODE == ROW_INIT_WARM ? b[idx] : r[idx];
                    w1[h][q] = Ap[idx];
                }
            }
    }
    // First chunk of the vector loads, issued before alpha is reduced (they
    // do not depend on it); further chunks (rows longer than 96) in the loop.
    constexpr int U = 6;
    RT v0[U], v1[U];
    float pv[U], xv[U];
    auto load = [&](int c0) {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = c0 + threadIdx.x + 256 * u;
            if (i < tot) {
                const long idx = base + i;
                if (MODE == ROW_INIT_COLD) {
                    v0[u] = b[idx];
                } else if (MODE == ROW_INIT_WARM) {
                    v0[u] = b[idx];
                    v1[u] = Ap[idx];
                } else {
                    v0[u] = r[idx];
                    v1[u] = Ap[idx];
                    if (MODE == ROW_LAST) {
                        pv[u] = p[idx];
                        xv[u] = x[idx];
                    }
                }
            }
        }
    };
    if constexpr (RP == 0) load(0);
    double alpha = 0.;
    if (MODE >= ROW_UPDATE) {
        const double pap = reduce_partials(SC.part_pap + (long)s * SC.maxpart, SC.npart_layer, sh, G.dot32);
        const double rz = SC.rz[(long)s * (SC.maxit + 2) + iter];
        double a = ddiv(rz, pap, G.dot32);
        const int stopped = stop[s];
        if (!isfinite(a) || stopped) {
            // lib/pcg.c:74-77: a non-finite alpha ends the solve; every
            // workgroup computes the same alpha, so all of them freeze x.
            a = 0;
            if (!stopped && threadIdx.x == 0 && bx == 0) stop[s] = 1;
        }
        alpha = a;
    }
    // x += alpha p of an UPDATE iteration is applied by the inverse row
    // kernel, which reads p anyway for p = z + beta p.
    if (MODE == ROW_UPDATE && bx == 0 && threadIdx.x == 0) SC.alpha[s] = alpha;
    FR *bufA = reinterpret_cast<FR *>(lds);
    double dot = 0;
    if constexpr (RP > 0 && MODE != ROW_LAST) {
        FT zr[RP];
#pragma unroll
        for (int q = 0; q < RP; q++) {
            RT rv[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const long idx = base + (long)(2 * fc + h) * n + lr + 32 * q;
                if (MODE == ROW_INIT_COLD) rv[h] = w0[h][q];
                else if (MODE == ROW_INIT_WARM) rv[h] = w0[h][q] - w1[h][q];
                else rv[h] = w0[h][q] - (RT)alpha * w1[h][q];
                if (fg < nf) {
                    r[idx] = rv[h];
                    if (MODE == ROW_INIT_COLD) x[idx] = 0.f;
                }
            }
            zr[q] = CT<FT>::make((FR)rv[0], (FR)rv[1]);
        }
        fft_fwd_l<RP, 32>(zr, lr, tw_of<FT>(G, false));
        const int kr = (int)(__builtin_bitreverse32((unsigned)lr) >> 27);
        if (fg < nf)
#pragma unroll
            for (int q = 0; q < RP; q++) lds[fc * n + q + RP * kr] = zr[q];
    } else {
    for (int c0 = 0; c0 < tot; c0 += 256 * U) {
        if (c0 > 0) load(c0);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = c0 + threadIdx.x + 256 * u;
            if (i < tot) {
                const int row = fdiv(i, ldf(G.dfnx)), col = i - row * n;
                const long idx = base + i;
                RT rv;
                if (MODE == ROW_INIT_COLD) {
                    rv = v0[u];
                    x[idx] = 0.f;
                } else if (MODE == ROW_INIT_WARM) {
                    rv = v0[u] - v1[u];
                } else {
                    if (MODE == ROW_LAST) x[idx] = xv[u] + (float)alpha * pv[u];
                    rv = v0[u] - (RT)alpha * v1[u];
                }
                r[idx] = rv;
                if (MODE == ROW_LAST) dot = dfma(rv, rv, dot, G.dot32);
                else bufA[2 * ((row >> 1) * n + col) + (row & 1)] = (FR)rv;
            }
        }
    }
    }
    if (MODE == ROW_LAST) {
        dot = block_sum(dot, sh, G.dot32);
        if (threadIdx.x == 0) SC.part_rr[(long)s * SC.maxpart + bx] = dot;
        return;
    }
    __syncthreads();
    FT *A = lds, *B = lds + ROWS / 2 * n;
    FT *res = RP > 0 ? A : fft_run<-1, FT, RM>(A, B, n, nf, G.nstx, G.fsx, twl);
    const int nh = n / 2 + 1;
    FT *S = spec + (long)s * G.nspec + G.soff[l];
    const FDiv dnf = ldf(G.dsmall[nf]);
    for (int i = threadIdx.x; i < nf * nh; i += blockDim.x) {
        const int k = fdiv(i, dnf), f = i - k * nf;
        const FT Z = res[f * n + k];
        const FT Zc = cconj(res[f * n + (k == 0 ? 0 : n - k)]);
        const FT sum = cadd(Z, Zc), dif = csub(Z, Zc);
        S[(long)k * nyl + row0 + 2 * f] = CT<FT>::make(FR(0.5) * sum.x, FR(0.5) * sum.y);
        S[(long)k * nyl + row0 + 2 * f + 1] = CT<FT>::make(FR(0.5) * dif.y, FR(-0.5) * dif.x);
    }
}

// ---------------------------------------------------------------------------
// Column kernel: forward column FFTs of one column group (all layers),
// per-frequency bs x bs block multiply (half spectrum, conjugate flags), then
// inverse column FFTs, for CS systems per workgroup (each block row of Minv
// and its positions are loaded once and applied to all CS systems).
// Everything in LDS: two buffers of CS * ncol * npsr * ny entries; the
// multiply writes the buffer the forward FFT did not end in, and the inverse
// FFT ping-pongs between it and the (then free) forward result.
// Global memory is touched in two rounds only: the block rows (into
// registers) and the spectrum columns at the start, the inverse columns at
// the end; every index comes from registers or LDS.
// ---------------------------------------------------------------------------

// One output row m of a block: y = sum_k M[m, k] v_k over the block's bins
// (k-th bin at LDS position pk[k], conjugated where flagged), written to C
// unless another block writes that bin, and its term of the spectral r.z:
// sc * w * Re(conj(R_m) y_m) with R the unscaled forward spectrum, y the
// scaled output (sc = fscale, or 1), w = 1 on the kx = 0 and nx/2 columns and
// 2 elsewhere (Parseval over the half spectrum); DC bins count zero when
// zero_dc (the blocks' DC columns are zero then, as if the input were).
// mk[k * ms]: row m's k-th coefficient.  KB: block size, 0 when runtime bs.
template <typename FT, int CS, int KB>
__device__ __forceinline__ void col_row(const float2 *mk, int ms, const int *pk, int pm, int bs, const FT *res,
                                        int per, int ns, float2 *C, double *dot, typename CT<FT>::R sc, int zd,
                                        int d32) {
    typedef typename CT<FT>::R FR;
    const int kb = KB > 0 ? KB : bs;
    FT y[CS];
#pragma unroll
    for (int si = 0; si < CS; si++) y[si] = CT<FT>::make(0, 0);
#pragma unroll
    for (int k = 0; k < kb; k++) {
        const int pe = pk[k];
        const int pos = pe & BP_POS;
        const FR cj = (pe & BP_CONJ) ? FR(-1) : FR(1);
        const float2 m = mk[k * ms];
        const FT mm = CT<FT>::make(m.x, m.y);
#pragma unroll
        for (int si = 0; si < CS; si++) {
            if (si < ns) {
                FT v = res[si * per + pos];
                v.y *= cj;
                y[si] = cadd(y[si], cmul(mm, v));
            }
        }
    }
    if (pm & BP_NOWRITE) return;
    const int pos = pm & BP_POS;
    const FR cj = (pm & BP_CONJ) ? FR(-1) : FR(1);
    const double wgt = (pm & BP_W1) ? 1.0 : 2.0;
    const bool skip = zd && (pm & BP_DC);
#pragma unroll
    for (int si = 0; si < CS; si++) {
        if (si < ns) {
            const FT yy = CT<FT>::make(sc * y[si].x, sc * cj * y[si].y);
            C[si * per + pos] = make_float2((float)yy.x, (float)yy.y);
            const FT rv = res[si * per + pos];
            if (!skip) dot[si] = dfma(wgt, dfma(rv.x, yy.x, dmul(rv.y, yy.y, d32), d32), dot[si], d32);
        }
    }
}

// block_sum of N values at once (the same order for each as block_sum):
// sh holds 5 * N doubles; 256-thread workgroups.
template <int N>
__device__ void block_sum_n(double *v, double *sh, int f32) {
#pragma unroll
    for (int n = 0; n < N; n++) {
        if (f32) {
            float f = (float)v[n];
            for (int o = 32; o > 0; o >>= 1) f += __shfl_down(f, o, 64);
            v[n] = f;
        } else {
            for (int o = 32; o > 0; o >>= 1) v[n] += __shfl_down(v[n], o, 64);
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int n = 0; n < N; n++) sh[n * 4 + wid] = v[n];
    __syncthreads();
    if (threadIdx.x < N) {
        double s = 0;
        float sf = 0.f;
        for (int i = 0; i < 4; i++) {
            s += sh[threadIdx.x * 4 + i];
            sf += (float)sh[threadIdx.x * 4 + i];
        }
        sh[4 * N + threadIdx.x] = f32 ? (double)sf : s;
    }
    __syncthreads();
#pragma unroll
    for (int n = 0; n < N; n++) v[n] = sh[4 * N + n];
    __syncthreads();
}

template <typename FT, int CS, int BS, int RM>
__global__ __launch_bounds__(256) void k_col(KARGS) {
    KJOB;
    const Scal SC = ldsc(PD.SC);
    const FT *specf = (const FT *)PD.specf;
    float2 *speci = PD.speci;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    __shared__ double sh[5 * CS];
    typedef typename CT<FT>::R FR;
    const int g = bx;
    const int s0 = J.s0 + bz * CS;
    const int ns = min(CS, J.s0 + J.nsys - s0);
    const int ny = G.fny, npsr = G.npsr;
    const int c0 = G.grp_col[g], c1 = G.grp_col[g + 1];
    const int ncg = c1 - c0;
    const int per = ncg * npsr * ny;  // entries of one system
    const int tot = ns * per;
    const int cap = CS * G.maxcol * npsr * ny;
    FT *A = reinterpret_cast<FT *>(lds_raw), *B = A + cap;
    FT *twf = B + cap;
    float2 *twi = reinterpret_cast<float2 *>(twf + ny);
    int *coff = reinterpret_cast<int *>(twi + ny);  // per (column, layer): spectrum offset of the column
    // Block rows of the group: with a compile-time block size, the first
    // RP * 256 rows are loaded into registers here, so that their latency
    // overlaps the spectrum loads and the forward FFT.
    const int bs = BS > 0 ? BS : G.bs;
    const int b0 = G.grp_blk[g], b1 = G.grp_blk[g + 1];
    const int nrow = (b1 - b0) * bs;
    constexpr int RP = BS > 0 ? 3 : 1;
    constexpr int BR = BS > 0 ? BS : 1;
    float2 mr[RP][BR];
    int pr[RP][BR], pmr[RP];
    if constexpr (BS > 0) {
#pragma unroll
        for (int u = 0; u < RP; u++) {
            // unconditional (row clamped; a group has at least one block)
            const int t = min((int)threadIdx.x + 256 * u, nrow - 1);
            const int q = t / BS, m = t - q * BS;
            const float2 *M = G.Minv + (long)(b0 + q) * BS * BS;
            const int *bp = G.blkpos + (long)(b0 + q) * BS;
#pragma unroll
            for (int k = 0; k < BS; k++) {
                mr[u][k] = M[m + k * BS];
                pr[u][k] = bp[k];
            }
            pmr[u] = bp[m];
        }
    }
    // (uniform loop: scalar loads, which do not wait behind the vector loads
    // above)
    for (int ci = 0; ci < ncg; ci++) {
        const int kx = ((const CONSTAS int *)G.colkx)[c0 + ci];
        for (int l = 0; l < npsr; l++)
            if (threadIdx.x == 0) coff[ci * npsr + l] = (int)G.soff[l] + kx * ny;
    }
    __syncthreads();
    // Twiddles and spectrum columns: U loads in flight per thread,
    // unconditional so that they issue back to back (past the end, the index
    // is clamped to the last entry, which is then loaded and stored again
    // with its own value); the twiddles go first and are stored with the
    // first batch.
    constexpr int U = 5;
    {
        const FT *twsf = tw_of<FT>(G, true);
        const float2 *twsi = G.twy;
        FT tf[(MAXN + 255) / 256];
        float2 ti[(MAXN + 255) / 256];
#pragma unroll
        for (int u = 0; u < (MAXN + 255) / 256; u++) {
            const int k = min((int)threadIdx.x + 256 * u, ny - 1);
            tf[u] = twsf[k];
            ti[u] = twsi[k];
        }
#pragma unroll
        for (int u = 0; u < (MAXN + 255) / 256; u++) {
            const int k = min((int)threadIdx.x + 256 * u, ny - 1);
            twf[k] = tf[u];
            twi[k] = ti[u];
        }
    }
    for (int i0 = 0; i0 < tot; i0 += 256 * U) {
        FT v[U];
        int ii[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = min(i0 + (int)threadIdx.x + 256 * u, tot - 1);
            const int si = CS == 1 ? 0 : (i >= per) + (i >= 2 * per) + (i >= 3 * per);
            const int j = i - si * per;
            const int f = fdiv(j, ldf(G.dfny)), ky = j - f * ny;
            ii[u] = i;
            v[u] = specf[(long)(s0 + si) * G.nspec + coff[f] + ky];
        }
#pragma unroll
        for (int u = 0; u < U; u++) A[ii[u]] = v[u];
    }
    __syncthreads();
    FT *res = fft_run<-1, FT, RM>(A, B, ny, ns * ncg * npsr, G.nsty, G.fsy, twf);
    float2 *C = reinterpret_cast<float2 *>(res == A ? B : A);
    float2 *D = reinterpret_cast<float2 *>(res);
    const FR sc = G.scale ? (FR)G.fscale : FR(1);
    const int zd = G.zero_dc, d32 = G.dot32;
    double dot[CS];
#pragma unroll
    for (int si = 0; si < CS; si++) dot[si] = 0;
    if constexpr (BS > 0) {
#pragma unroll
        for (int u = 0; u < RP; u++) {
            const int t = threadIdx.x + 256 * u;
            if (t < nrow) col_row<FT, CS, BS>(mr[u], 1, pr[u], pmr[u], BS, res, per, ns, C, dot, sc, zd, d32);
        }
    }
    for (int t = threadIdx.x + 256 * (BS > 0 ? RP : 0); t < nrow; t += blockDim.x) {
        const int q = t / bs, m = t - q * bs;
        const float2 *M = G.Minv + (long)(b0 + q) * bs * bs;
        const int *bp = G.blkpos + (long)(b0 + q) * bs;
        col_row<FT, CS, BS>(M + m, bs, bp, bp[m], bs, res, per, ns, C, dot, sc, zd, d32);
    }
    block_sum_n<CS>(dot, sh, d32);  // its barriers also order C before the inverse FFT
    if (threadIdx.x == 0)
#pragma unroll
        for (int si = 0; si < CS; si++)
            if (si < ns)
                SC.part_rz[(long)(s0 + si) * SC.maxpart + g] =
                    d32 ? (double)((float)dot[si] * (float)sc) : dot[si] * (double)sc;
    float2 *res2 = fft_run<1, float2, RM>(C, D, ny, ns * ncg * npsr, G.nsty, G.fsy, twi);
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        const int si = CS == 1 ? 0 : (i >= per) + (i >= 2 * per) + (i >= 3 * per);
        const int j = i - si * per;
        const int f = fdiv(j, ldf(G.dfny)), ky = j - f * ny;
        speci[(long)(s0 + si) * G.nspec + coff[f] + ky] = res2[i];
    }
}

// ---------------------------------------------------------------------------
// Column kernel with the column FFTs in registers, for FDPCG blocks that
// each couple one frequency (kx, ky) of the half spectrum across the
// bs = npsr layers (no aliasing partners: every column is a group of its
// own).  A task is one (system, column) on bs x L threads: L lanes per layer
// hold its ny = L * P points (lane l: ky = l + L p).  Each lane group runs
// the forward FFT of its layer across its lanes (then lane l holds the bins
// ky = k1 + P bitrev(l)), puts them in LDS, and after one barrier each lane
// reads the same bins of the other layers, applies its row of each bin's
// block, and runs the inverse FFT of its layer, again in registers.  The
// block rows are stored per (column, k1, entry) with the L lanes innermost
// (Mreg), conjugated where the block's frequency is the conjugate of its
// half-spectrum bin and with the DC columns zeroed; they are loaded at the
// start, under the spectrum loads and the forward FFT.  Per bin, as the LDS
// kernel: y_m = sum_k M[m + k bs] x_k in the spectrum type (here with fused
// multiply-adds), scaled, rounded to FP32 for the inverse; r.z adds
// w * Re(conj(x_m) y_m) (w = 1 on the kx = 0 and Nyquist columns, else 2;
// DC skipped), reduced per lane group, then over the layers in order, into
// part_rz[s][column].  Workgroup: T = max(1, 256 / (bs L)) tasks, systems
// bx * T + t of column by; blockDim.x = T bs L rounded up to 64 (idle lanes
// repeat the last lane group and write nothing).  LT: L at compile time
// (32: unrolled stages, exchanges without LDS addresses), or 0 (runtime);
// BS: the block size at compile time, or 0 (runtime, at most COLR_MAXBS).
// ---------------------------------------------------------------------------
constexpr int COLR_MAXBS = 8;
__host__ __device__ inline int colr_tasks(int bs, int L) { return bs * L >= 256 ? 1 : 256 / (bs * L); }

template <typename FT, int P, int LT, int BS>
__global__ __launch_bounds__(512) void k_colr(KARGS) {
    KJOB;
    const Scal SC = ldsc(PD.SC);
    typedef typename CT<FT>::R FR;
    constexpr int NB = BS ? BS : COLR_MAXBS;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const FT *specf = (const FT *)PD.specf;
    const int L = LT ? LT : G.shL, ny = G.fny, bs = BS ? BS : G.bs;
    const int T = colr_tasks(bs, L), per = bs * L, lb = 31 - __builtin_clz(L);
    const int t = min((int)threadIdx.x, T * per - 1);  // idle lanes: the last lane group again
    const int ti = t / per, q = t - ti * per, m = q >> lb, l = q & (L - 1);
    const bool lane_live = (int)threadIdx.x < T * per;
    const int si = bx * T + ti;
    const bool live = lane_live && si < J.nsys;
    const int s = J.s0 + min(si, J.nsys - 1);
    const int c = by, kx = ((const CONSTAS int *)G.colkx)[c];
    const long sidx = (long)s * G.nspec + G.soff[m] + (long)kx * ny + l;
    FT x[P];
#pragma unroll
    for (int p = 0; p < P; p++) x[p] = specf[sidx + L * p];
    // row m of the blocks of this lane's bins (a uniform base per entry,
    // the lane's offset m L + l)
    const float2 *Mcol = G.Mreg + (long)c * P * bs * bs * L;
    const unsigned moff = (unsigned)(m * L + l);
    float2 mrow[P][NB];
#pragma unroll
    for (int k1 = 0; k1 < P; k1++)
#pragma unroll
        for (int k = 0; k < NB; k++)
            if (k < bs) mrow[k1][k] = (Mcol + (k1 * bs * bs + k * bs) * L)[moff];
    if constexpr (LT > 0) fft_fwd_l<P, LT>(x, l, tw_of<FT>(G, true));
    else fft_regs_fwd<P>(x, L, l, tw_of<FT>(G, true));
    // bins of every layer of the task: [task][layer][k1][lane]
    FT *X = reinterpret_cast<FT *>(lds_raw);
    double *dsum = reinterpret_cast<double *>(X + (long)T * per * P);
    FT *Xt = X + (long)ti * per * P;
    if (lane_live)
#pragma unroll
        for (int k1 = 0; k1 < P; k1++) Xt[(m * P + k1) * L + l] = x[k1];
    __syncthreads();
    const int kr = L > 1 ? (int)(__builtin_bitreverse32((unsigned)l) >> (32 - lb)) : 0;
    const FR sc = G.scale ? (FR)G.fscale : FR(1);
    const int zd = G.zero_dc, d32 = G.dot32;
    const double wgt = (kx == 0 || 2 * kx == G.fnx) ? 1.0 : 2.0;
    float2 z[P];
    double dot = 0;
#pragma unroll
    for (int k1 = 0; k1 < P; k1++) {
        FR yr = 0, yi = 0;
#pragma unroll
        for (int k = 0; k < NB; k++) {
            if (k < bs) {
                const FR mr = mrow[k1][k].x, mi = mrow[k1][k].y;
                const FT v = Xt[(k * P + k1) * L + l];
                yr = fma(mr, v.x, yr);
                yr = fma(-mi, v.y, yr);
                yi = fma(mr, v.y, yi);
                yi = fma(mi, v.x, yi);
            }
        }
        const FT y = CT<FT>::make(yr, yi);
        const FT yy = CT<FT>::make(sc * y.x, sc * y.y);
        z[k1] = make_float2((float)yy.x, (float)yy.y);
        if (!(zd && kx == 0 && k1 == 0 && kr == 0))
            dot = dfma(wgt, dfma(x[k1].x, yy.x, dmul(x[k1].y, yy.y, d32), d32), dot, d32);
    }
    // r.z: over the lane group, then over the layers in order
    if (d32) {
        float f = (float)dot;
        if constexpr (LT > 0) f = lane_sum<LT>(f);
        else
            for (int h = L >> 1; h >= 1; h >>= 1) f += __shfl_xor(f, h, 64);
        dot = f;
    } else {
        if constexpr (LT > 0) dot = lane_sum<LT>(dot);
        else
            for (int h = L >> 1; h >= 1; h >>= 1) dot += __shfl_xor(dot, h, 64);
    }
    if (lane_live && l == 0) dsum[ti * bs + m] = dot;
    if constexpr (LT > 0) fft_inv_l<P, LT>(z, l, G.twy);
    else fft_regs_inv<P>(z, L, l, G.twy);
    if (live) {
        float2 *dst = PD.speci + sidx;
#pragma unroll
        for (int p = 0; p < P; p++) dst[L * p] = z[p];
    }
    __syncthreads();
    if ((int)threadIdx.x < T && bx * T + (int)threadIdx.x < J.nsys) {
        const double *d = dsum + threadIdx.x * bs;
        double v;
        if (d32) {
            float f = 0.f;
            for (int k = 0; k < bs; k++) f += (float)d[k];
            v = (double)(f * (float)sc);
        } else {
            v = 0;
            for (int k = 0; k < bs; k++) v += d[k];
            v *= (double)sc;
        }
        SC.part_rz[(long)(J.s0 + bx * T + threadIdx.x) * SC.maxpart + c] = v;
    }
}

// ---------------------------------------------------------------------------
// Inverse row kernel: complex-to-real row FFT -> z (the imaginary parts of
// the DC and Nyquist bins are ignored, as C2R does), then by mode
//   IROW_Z       out = z                                  (P alone)
//   IROW_FIRST   p = z; rz[slot] = r.z                    (start / restart)
//   IROW_UPDATE  beta = r.z / rz[slot]; rz[slot+1] = r.z;
//                x += alpha p; p = z + beta p             (lib/pcg.c:703,714)
// r.z is reduced from the column kernel's spectral partials (fixed order,
// the same value in every workgroup).
// ---------------------------------------------------------------------------
enum { IROW_Z = 0, IROW_FIRST = 1, IROW_UPDATE = 2 };

template <int MODE, int RM, int RP>
__global__ __launch_bounds__(256) void k_irow(KARGS, int slot, int osel) {
    KJOB;
    const Scal SC = ldsc(PD.SC);
    const float2 *spec = PD.speci;
    float *x = PD.x;
    float *out = vsel(PD, osel);
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float2 *lds = reinterpret_cast<float2 *>(lds_raw);
    __shared__ double sh[8];
    const int s = J.s0 + bz;
    const int2 rb = G.rowblk[bx];
    const int l = rb.x, row0 = rb.y;
    const int n = G.fnx, nyl = G.xny[l];
    const int nrows = min(ROWS, nyl - row0);
    const int nf = nrows / 2;
    const float2 *S = spec + (long)s * G.nspec + G.soff[l];
    float2 *A = lds, *B = lds + ROWS / 2 * n;
    float2 *twl = lds + ROWS * n;
    if (RP == 0) load_tw(twl, G.twx, n);
    // Loads first (spectrum, and x, p for the update), then the r.z
    // reduction while they are in flight.
    const int tots = nf * n;
    const FDiv dnf = ldf(G.dsmall[nf]);
    constexpr int US = 4;
    float2 va[US], vb[US];
    auto load_spec = [&](int c0) {
#pragma unroll
        for (int u = 0; u < US; u++) {
            const int i = c0 + threadIdx.x + 256 * u;
            if (i < tots) {
                const int k = fdiv(i, dnf), f = i - k * nf;
                const int ya = row0 + 2 * f;
                const long o = (long)(k <= n / 2 ? k : n - k) * nyl + ya;
                va[u] = S[o];
                vb[u] = S[o + 1];
            }
        }
    };
    auto store_spec = [&](int c0) {
#pragma unroll
        for (int u = 0; u < US; u++) {
            const int i = c0 + threadIdx.x + 256 * u;
            if (i < tots) {
                const int k = fdiv(i, dnf), f = i - k * nf;
                float2 a = va[u], b = vb[u];
                if (k > n / 2) {
                    a = cconj(a);
                    b = cconj(b);
                } else if (k == 0 || k == n / 2) {
                    a.y = 0.f;
                    b.y = 0.f;
                }
                A[f * n + k] = make_float2(a.x - b.y, a.y + b.x);
            }
        }
    };
    const long base = (long)s * G.nxtot + G.xoff[l] + (long)row0 * n;
    const int tot = nrows * n;
    constexpr int U = 6;
    float pv[U], xv[U];
    auto load_xp = [&](int c0) {
        if (MODE != IROW_UPDATE) return;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = c0 + threadIdx.x + 256 * u;
            if (i < tot) {
                pv[u] = out[base + i];
                xv[u] = x[base + i];
            }
        }
    };
    load_spec(0);
    // register path (RP > 0): lane group f takes rows 2f, 2f + 1, lane lr
    // columns lr + 32 p; its x and p are loaded here
    constexpr int RPL = RP > 0 ? RP : 1;
    const int fg = threadIdx.x >> 5, lr = threadIdx.x & 31;
    const int fc = min(fg, nf - 1);
    float px[2][RPL], xx[2][RPL];
    if constexpr (RP > 0) {
        if (MODE == IROW_UPDATE)
#pragma unroll
            for (int h = 0; h < 2; h++)
#pragma unroll
                for (int q = 0; q < RP; q++) {
                    const long idx = base + (long)(2 * fc + h) * n + lr + 32 * q;
                    px[h][q] = out[idx];
                    xx[h][q] = x[idx];
                }
    } else {
        load_xp(0);
    }
    double beta = 0, alpha = 0;
    if (MODE != IROW_Z) {
        const double rz = reduce_partials(SC.part_rz + (long)s * SC.maxpart, SC.npart_col, sh, G.dot32);
        double *hist = SC.rz + (long)s * (SC.maxit + 2);
        if (MODE == IROW_UPDATE) {
            beta = ddiv(rz, hist[slot], G.dot32);
            if (!isfinite(beta)) beta = 0;
            alpha = SC.alpha[s];
        }
        if (bx == 0 && threadIdx.x == 0) hist[MODE == IROW_FIRST ? slot : slot + 1] = rz;
    }
    for (int c0 = 0; c0 < tots; c0 += 256 * US) {
        if (c0 > 0) load_spec(c0);
        store_spec(c0);
    }
    __syncthreads();
    const float sc = G.scale ? G.fscale : 1.f;
    const float fa = (float)alpha, fb = (float)beta;
    if constexpr (RP > 0) {
        const int kr = (int)(__builtin_bitreverse32((unsigned)lr) >> 27);
        float2 y[RP];
#pragma unroll
        for (int q = 0; q < RP; q++) y[q] = A[fc * n + q + RP * kr];
        fft_inv_l<RP, 32>(y, lr, G.twx);
        if (fg < nf)
#pragma unroll
            for (int q = 0; q < RP; q++)
#pragma unroll
                for (int h = 0; h < 2; h++) {
                    const long idx = base + (long)(2 * fc + h) * n + lr + 32 * q;
                    const float zv = (h ? y[q].y : y[q].x) * sc;
                    if (MODE == IROW_UPDATE) {
                        x[idx] = xx[h][q] + fa * px[h][q];
                        out[idx] = zv + fb * px[h][q];
                    } else {
                        out[idx] = zv;
                    }
                }
        return;
    }
    float2 *res = fft_run<1, float2, RM>(A, B, n, nf, G.nstx, G.fsx, twl);
    for (int c0 = 0; c0 < tot; c0 += 256 * U) {
        if (c0 > 0) load_xp(c0);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int i = c0 + threadIdx.x + 256 * u;
            if (i < tot) {
                const int row = fdiv(i, ldf(G.dfnx)), col = i - row * n;
                const float2 v = res[(row >> 1) * n + col];
                const float zv = ((row & 1) ? v.y : v.x) * sc;
                if (MODE == IROW_UPDATE) {
                    x[base + i] = xv[u] + fa * pv[u];
                    out[base + i] = zv + fb * pv[u];
                } else {
                    out[base + i] = zv;
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Plain CG (tomo.precond=0: lib/pcg.c:28-113 with Mmul = NULL, so z = r),
// over the row blocks of the layers (rowblk: layer, first row; ROWS rows of
// that layer's width), so the r.r partials have the row kernels' count:
//   k_cgv  INIT_COLD  r = b, x = 0                      r.r partial
//          INIT_WARM  r = b - A x (Ap holds A x)        r.r partial
//          UPDATE     alpha = rz[iter] / p.Ap;
//                     x += alpha p; r -= alpha Ap       r.r partial
//          LAST       the same (r.r of the final residual)
//   k_cgp  FIRST      rz[slot] = r.r; p = r             (start / restart)
//          UPDATE     beta = r.r / rz[slot]; rz[slot+1] = r.r; p = r + beta p
// alpha, beta and the sums are FP64 and reduced in fixed order (the same
// value in every workgroup); a non-finite alpha stops the system as k_row
// does (lib/pcg.c:74-77).
// ---------------------------------------------------------------------------
enum { CGV_INIT_COLD = 0, CGV_INIT_WARM = 1, CGV_UPDATE = 2, CGV_LAST = 3 };
template <int MODE, typename RT>
__global__ __launch_bounds__(256) void k_cgv(KARGS, int iter) {
    KJOB;
    const Scal SC = ldsc(PD.SC);
    float *x = PD.x;
    const float *p = PD.p;
    RT *r = (RT *)PD.r;
    const RT *b = (const RT *)PD.b, *Ap = (const RT *)PD.Ap;
    int *stop = PD.stop;
    __shared__ double sh[8];
    const int s = J.s0 + bz;
    const int2 rb = G.rowblk[bx];
    const int l = rb.x, row0 = rb.y;
    const int n = G.xnx[l];
    const int tot = min(ROWS, G.xny[l] - row0) * n;
    const long base = (long)s * G.nxtot + G.xoff[l] + (long)row0 * n;
    double alpha = 0.;
    if (MODE >= CGV_UPDATE) {
        const double pap = reduce_partials(SC.part_pap + (long)s * SC.maxpart, SC.npart_layer, sh, G.dot32);
        const double rz = SC.rz[(long)s * (SC.maxit + 2) + iter];
        double a = ddiv(rz, pap, G.dot32);
        const int stopped = stop[s];
        if (!isfinite(a) || stopped) {
            a = 0;
            if (!stopped && threadIdx.x == 0 && bx == 0) stop[s] = 1;
        }
        alpha = a;
    }
    double dot = 0;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        const long idx = base + i;
        RT rv;
        if (MODE == CGV_INIT_COLD) {
            rv = b[idx];
            x[idx] = 0.f;
        } else if (MODE == CGV_INIT_WARM) {
            rv = b[idx] - Ap[idx];
        } else {
            const float pv = p[idx];
            x[idx] = x[idx] + (float)alpha * pv;
            rv = r[idx] - (RT)alpha * Ap[idx];
        }
        r[idx] = rv;
        dot = dfma(rv, rv, dot, G.dot32);
    }
    dot = block_sum(dot, sh, G.dot32);
    if (threadIdx.x == 0) SC.part_rr[(long)s * SC.maxpart + bx] = dot;
}
enum { CGP_FIRST = 0, CGP_UPDATE = 1 };
template <int MODE, typename RT>
__global__ __launch_bounds__(256) void k_cgp(KARGS, int slot) {
    KJOB;
    const Scal SC = ldsc(PD.SC);
    float *p = PD.p;
    const RT *r = (const RT *)PD.r;
    __shared__ double sh[8];
    const int s = J.s0 + bz;
    const int2 rb = G.rowblk[bx];
    const int l = rb.x, row0 = rb.y;
    const int n = G.xnx[l];
    const int tot = min(ROWS, G.xny[l] - row0) * n;
    const long base = (long)s * G.nxtot + G.xoff[l] + (long)row0 * n;
    const double rr = reduce_partials(SC.part_rr + (long)s * SC.maxpart, SC.npart_row, sh, G.dot32);
    double *hist = SC.rz + (long)s * (SC.maxit + 2);
    double beta = 0;
    if (MODE == CGP_UPDATE) {
        beta = ddiv(rr, hist[slot], G.dot32);
        if (!isfinite(beta)) beta = 0;
    }
    if (bx == 0 && threadIdx.x == 0) hist[MODE == CGP_FIRST ? slot : slot + 1] = rr;
    const float fb = (float)beta;
    for (int i = threadIdx.x; i < tot; i += blockDim.x) {
        const long idx = base + i;
        const float rv = (float)r[idx];
        p[idx] = MODE == CGP_FIRST ? rv : rv + fb * p[idx];
    }
}

// Residual sqrt(r.r / b.b) per system (NaN when the solve stopped early).
__global__ __launch_bounds__(256) void k_final(KARGS) {
    KJOB;
    const Scal SC = ldsc(PD.SC);
    const int *stop = PD.stop;
    const int npart_rr = G.nrowblk;
    __shared__ double sh[8];
    const int s = J.s0 + bz;
    const double rr = reduce_partials(SC.part_rr + (long)s * SC.maxpart, npart_rr, sh);
    const double bb = reduce_partials(SC.part_bb + (long)s * SC.maxpart, SC.npart_layer, sh);
    if (threadIdx.x == 0) SC.res[s] = stop[s] ? NAN : sqrt(rr / bb);
}

// Copy / convert helpers.
// Request-loop staging (one copy each way per batch instead of one per
// system): SB workgroups per job; jobs that are off exit at entry.
constexpr int SB = 16;
__global__ __launch_bounds__(256) void k_stage_in(const Job *__restrict__ jobs, const StageOff *__restrict__ off,
                                                  const float *__restrict__ src) {
    const int j = blockIdx.x / SB, part = blockIdx.x - j * SB;
    const Job J = jobs[j];
    if (!J.on) return;
    CPD &PD = *(const CONSTAS PlanDev *)J.pd;
    const long ng = PD.G.ngtot, n = (long)J.nsys * ng;
    float *dst = PD.g + (long)J.s0 * ng;
    const float *s = src + off[j].g;
    for (long i = (long)part * 256 + threadIdx.x; i < n; i += (long)SB * 256) dst[i] = s[i];
}
__global__ __launch_bounds__(256) void k_stage_out(const Job *__restrict__ jobs, const StageOff *__restrict__ off,
                                                   float *__restrict__ dx, double *__restrict__ dres) {
    const int j = blockIdx.x / SB, part = blockIdx.x - j * SB;
    const Job J = jobs[j];
    if (!J.on) return;
    CPD &PD = *(const CONSTAS PlanDev *)J.pd;
    const long nx = PD.G.nxtot, n = (long)J.nsys * nx;
    const float *s = PD.x + (long)J.s0 * nx;
    float *d = dx + off[j].x;
    for (long i = (long)part * 256 + threadIdx.x; i < n; i += (long)SB * 256) d[i] = s[i];
    if (part == 0)
        for (int k = threadIdx.x; k < J.nsys; k += 256) dres[off[j].r + k] = PD.SC.res[J.s0 + k];
}

__global__ void k_copy(const float *a, float *b, long n) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}
// Unit vectors: slot k (blockIdx.y) of n slots of len values each is zero
// but for 1 at idx[k] (all zero when idx is null).
__global__ void k_unit(float *v, long len, const int *idx) {
    const int k = blockIdx.y;
    const long one = idx ? idx[k] : -1;
    float *vk = v + (long)k * len;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < len; i += (long)gridDim.x * blockDim.x)
        vk[i] = i == one ? 1.f : 0.f;
}
__global__ void k_zero_stop(KARGS) {
    KJOB;
    for (int i = threadIdx.x; i < J.nsys; i += blockDim.x) PD.stop[J.s0 + i] = 0;
}

// ---------------------------------------------------------------------------
// launch wrappers: one 1-D grid of t.nwg workgroups, each finding its job and
// block in t.wgi.  Template choices (precision, radix set, block size,
// systems per workgroup) are uniform over a launch.
// ---------------------------------------------------------------------------
#define KT t.jobs, t.wgi
void launch_ttr(const LaunchTab &t, hipStream_t st) {
    hipLaunchKernelGGL(k_ttr, dim3(t.nwg), dim3(256), 0, st, KT);
}
template <typename RT>
static void launch_wfs_rt(const LaunchTab &t, bool rhs, int matched, int pos, int cs, int xsel, hipStream_t st) {
    const dim3 grid(t.nwg), blk(256);
    if (rhs) {
        if (pos == 1) hipLaunchKernelGGL((k_wfs<1, true, RT>), grid, blk, 0, st, KT, xsel);
        else hipLaunchKernelGGL((k_wfs<2, true, RT>), grid, blk, 0, st, KT, xsel);
    } else if (!matched) {
        if (pos == 1) hipLaunchKernelGGL((k_wfs<1, false, RT>), grid, blk, 0, st, KT, xsel);
        else hipLaunchKernelGGL((k_wfs<2, false, RT>), grid, blk, 0, st, KT, xsel);
    } else if (pos == 1) {
        if (cs == 4) hipLaunchKernelGGL((k_wfs_m<1, RT, 4>), grid, blk, 0, st, KT, xsel);
        else if (cs == 2) hipLaunchKernelGGL((k_wfs_m<1, RT, 2>), grid, blk, 0, st, KT, xsel);
        else hipLaunchKernelGGL((k_wfs_m<1, RT, 1>), grid, blk, 0, st, KT, xsel);
    } else {
        if (cs >= 2) hipLaunchKernelGGL((k_wfs_m<2, RT, 2>), grid, blk, 0, st, KT, xsel);
        else hipLaunchKernelGGL((k_wfs_m<2, RT, 1>), grid, blk, 0, st, KT, xsel);
    }
}
void launch_wfs(const LaunchTab &t, bool rhs, int matched, int pos, int cs, int rt64, int xsel, hipStream_t st) {
    if (rt64) launch_wfs_rt<double>(t, rhs, matched, pos, cs, xsel, st);
    else launch_wfs_rt<float>(t, rhs, matched, pos, cs, xsel, st);
}
template <bool LHS, typename RT>
static void launch_layer_rt(const LaunchTab &t, int matched, int cs, int psel_, int osel, int partsel,
                            hipStream_t st) {
    const dim3 grid(t.nwg), blk(256);
    if (!matched) hipLaunchKernelGGL((k_layer<LHS, RT>), grid, blk, 0, st, KT, psel_, osel, partsel);
    else if (matched == 2) {  // 16 x 64 tiles
        const dim3 b8(128);
        if (cs == 4) hipLaunchKernelGGL((k_layer_m8<LHS, RT, 4, LTY>), grid, b8, 0, st, KT, psel_, osel, partsel);
        else if (cs == 2) hipLaunchKernelGGL((k_layer_m8<LHS, RT, 2, LTY>), grid, b8, 0, st, KT, psel_, osel, partsel);
        else hipLaunchKernelGGL((k_layer_m8<LHS, RT, 1, LTY>), grid, b8, 0, st, KT, psel_, osel, partsel);
    } else if (matched == 3) {  // 16 x 32 tiles
        const dim3 b8(128);
        if (cs == 4) hipLaunchKernelGGL((k_layer_m8<LHS, RT, 4, LTY2>), grid, b8, 0, st, KT, psel_, osel, partsel);
        else if (cs == 2) hipLaunchKernelGGL((k_layer_m8<LHS, RT, 2, LTY2>), grid, b8, 0, st, KT, psel_, osel, partsel);
        else hipLaunchKernelGGL((k_layer_m8<LHS, RT, 1, LTY2>), grid, b8, 0, st, KT, psel_, osel, partsel);
    }
    else if (cs == 4) hipLaunchKernelGGL((k_layer_m<LHS, RT, 4>), grid, blk, 0, st, KT, psel_, osel, partsel);
    else if (cs == 2) hipLaunchKernelGGL((k_layer_m<LHS, RT, 2>), grid, blk, 0, st, KT, psel_, osel, partsel);
    else hipLaunchKernelGGL((k_layer_m<LHS, RT, 1>), grid, blk, 0, st, KT, psel_, osel, partsel);
}
void launch_layer(const LaunchTab &t, bool lhs, int matched, int cs, int rt64, int psel_, int osel, int partsel,
                  hipStream_t st) {
    if (rt64) {
        if (lhs) launch_layer_rt<true, double>(t, matched, cs, psel_, osel, partsel, st);
        else launch_layer_rt<false, double>(t, matched, cs, psel_, osel, partsel, st);
    } else {
        if (lhs) launch_layer_rt<true, float>(t, matched, cs, psel_, osel, partsel, st);
        else launch_layer_rt<false, float>(t, matched, cs, psel_, osel, partsel, st);
    }
}
// LDS of the row kernels: ROWS/2 complex rows twice (ping-pong) + twiddles.
size_t row_lds(int fnx, int f64) { return (f64 ? sizeof(double2) : sizeof(float2)) * (ROWS + 1) * (size_t)fnx; }
template <typename RT, typename FT, int RM, int RP>
static void launch_row_r(const LaunchTab &t, int mode, int iter, int xsel, size_t lds, hipStream_t st) {
    const dim3 grid(t.nwg), blk(256);
    switch (mode) {
        case ROW_INIT_COLD:
            hipLaunchKernelGGL((k_row<ROW_INIT_COLD, RT, FT, RM, RP>), grid, blk, lds, st, KT, iter, xsel);
            break;
        case ROW_INIT_WARM:
            hipLaunchKernelGGL((k_row<ROW_INIT_WARM, RT, FT, RM, RP>), grid, blk, lds, st, KT, iter, xsel);
            break;
        case ROW_UPDATE:
            hipLaunchKernelGGL((k_row<ROW_UPDATE, RT, FT, RM, RP>), grid, blk, lds, st, KT, iter, xsel);
            break;
        default: hipLaunchKernelGGL((k_row<ROW_LAST, RT, FT, RM, 0>), grid, blk, lds, st, KT, iter, xsel); break;
    }
}
template <typename RT, typename FT>
static void launch_row_t(const LaunchTab &t, int mode, int rm, int rp, int iter, int xsel, size_t lds,
                         hipStream_t st) {
    if (rm > 4) launch_row_r<RT, FT, 7, 0>(t, mode, iter, xsel, lds, st);
    else if (rp == 3) launch_row_r<RT, FT, 3, 3>(t, mode, iter, xsel, lds, st);
    else if (rp == 2) launch_row_r<RT, FT, 3, 2>(t, mode, iter, xsel, lds, st);
    else launch_row_r<RT, FT, 3, 0>(t, mode, iter, xsel, lds, st);
}
void launch_row(const LaunchTab &t, int mode, int rt64, int spec64, int rm, int rp, int iter, int xsel, size_t lds,
                hipStream_t st) {
    if (rt64 && spec64) launch_row_t<double, double2>(t, mode, rm, rp, iter, xsel, lds, st);
    else if (rt64) launch_row_t<double, float2>(t, mode, rm, rp, iter, xsel, lds, st);
    else launch_row_t<float, float2>(t, mode, rm, rp, iter, xsel, lds, st);
}
// LDS of the column kernel: two buffers of cs systems' column groups + the
// forward and inverse twiddles.
size_t col_lds(int maxcol, int npsr, int fny, int f64, int cs) {
    return (f64 ? sizeof(double2) : sizeof(float2)) * (2 * (size_t)cs * maxcol * npsr * fny + fny) +
           sizeof(float2) * fny + sizeof(int) * (size_t)maxcol * npsr;
}
template <typename FT, int CS, int RM>
static void launch_col_cs(const LaunchTab &t, int bs, size_t lds, hipStream_t st) {
    const dim3 grid(t.nwg), blk(256);
    switch (bs) {
        case 2: hipLaunchKernelGGL((k_col<FT, CS, 2, RM>), grid, blk, lds, st, KT); break;
        case 3: hipLaunchKernelGGL((k_col<FT, CS, 3, RM>), grid, blk, lds, st, KT); break;
        case 4: hipLaunchKernelGGL((k_col<FT, CS, 4, RM>), grid, blk, lds, st, KT); break;
        case 6: hipLaunchKernelGGL((k_col<FT, CS, 6, RM>), grid, blk, lds, st, KT); break;
        case 8: hipLaunchKernelGGL((k_col<FT, CS, 8, RM>), grid, blk, lds, st, KT); break;
        default: hipLaunchKernelGGL((k_col<FT, CS, 0, RM>), grid, blk, lds, st, KT); break;
    }
}
// The instantiation must be the systems per workgroup the grid and the LDS
// size were built for (1, 2 or 4).
template <typename FT, int RM>
static void launch_col_r(const LaunchTab &t, int cs, int bs, size_t lds, hipStream_t st) {
    if (cs == 4) launch_col_cs<FT, 4, RM>(t, bs, lds, st);
    else if (cs == 2) launch_col_cs<FT, 2, RM>(t, bs, lds, st);
    else launch_col_cs<FT, 1, RM>(t, bs, lds, st);
}
template <typename FT, int CS, int RM>
static void col_allow_bs(int bs, size_t lds) {
    const void *f;
    switch (bs) {
        case 2: f = (const void *)k_col<FT, CS, 2, RM>; break;
        case 3: f = (const void *)k_col<FT, CS, 3, RM>; break;
        case 4: f = (const void *)k_col<FT, CS, 4, RM>; break;
        case 6: f = (const void *)k_col<FT, CS, 6, RM>; break;
        case 8: f = (const void *)k_col<FT, CS, 8, RM>; break;
        default: f = (const void *)k_col<FT, CS, 0, RM>; break;
    }
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}
template <typename FT, int RM>
static void col_allow_r(int cs, int bs, size_t lds) {
    if (cs == 4) col_allow_bs<FT, 4, RM>(bs, lds);
    else if (cs == 2) col_allow_bs<FT, 2, RM>(bs, lds);
    else col_allow_bs<FT, 1, RM>(bs, lds);
}
void col_allow_lds(int spec64, int cs, int bs, int rm, size_t lds) {
    if (spec64) {
        if (rm > 4) col_allow_r<double2, 7>(cs, bs, lds);
        else col_allow_r<double2, 3>(cs, bs, lds);
    } else {
        if (rm > 4) col_allow_r<float2, 7>(cs, bs, lds);
        else col_allow_r<float2, 3>(cs, bs, lds);
    }
}
void launch_col(const LaunchTab &t, int spec64, int cs, int bs, int rm, size_t lds, hipStream_t st) {
    if (spec64) {
        if (rm > 4) launch_col_r<double2, 7>(t, cs, bs, lds, st);
        else launch_col_r<double2, 3>(t, cs, bs, lds, st);
    } else {
        if (rm > 4) launch_col_r<float2, 7>(t, cs, bs, lds, st);
        else launch_col_r<float2, 3>(t, cs, bs, lds, st);
    }
}
template <typename FT>
static void launch_colr_t(const LaunchTab &t, int P, int L, int bs, int nthr, size_t lds, hipStream_t st) {
    const dim3 grid(t.nwg), blk(nthr);
    if (L == 32 && P == 3 && bs == 6) hipLaunchKernelGGL((k_colr<FT, 3, 32, 6>), grid, blk, lds, st, KT);
    else if (L == 32 && P == 3) hipLaunchKernelGGL((k_colr<FT, 3, 32, 0>), grid, blk, lds, st, KT);
    else if (L == 32 && P == 2) hipLaunchKernelGGL((k_colr<FT, 2, 32, 0>), grid, blk, lds, st, KT);
    else switch (P) {
        case 2: hipLaunchKernelGGL((k_colr<FT, 2, 0, 0>), grid, blk, lds, st, KT); break;
        case 3: hipLaunchKernelGGL((k_colr<FT, 3, 0, 0>), grid, blk, lds, st, KT); break;
        case 4: hipLaunchKernelGGL((k_colr<FT, 4, 0, 0>), grid, blk, lds, st, KT); break;
        case 5: hipLaunchKernelGGL((k_colr<FT, 5, 0, 0>), grid, blk, lds, st, KT); break;
        default: hipLaunchKernelGGL((k_colr<FT, 7, 0, 0>), grid, blk, lds, st, KT); break;
    }
}
bool colr_supported(int P, int bs) { return (P == 2 || P == 3 || P == 4 || P == 5 || P == 7) && bs >= 1 && bs <= COLR_MAXBS; }
int colr_shape(int bs, int L, int *tasks) {
    const int T = colr_tasks(bs, L);
    if (tasks) *tasks = T;
    return (T * bs * L + 63) / 64 * 64;
}
size_t colr_lds(int bs, int L, int P, int f64) {
    const int T = colr_tasks(bs, L);
    return (size_t)T * bs * L * P * (f64 ? sizeof(double2) : sizeof(float2)) + sizeof(double) * (size_t)T * bs;
}
void launch_colr(const LaunchTab &t, int spec64, int P, int L, int bs, int nthr, size_t lds, hipStream_t st) {
    if (spec64) launch_colr_t<double2>(t, P, L, bs, nthr, lds, st);
    else launch_colr_t<float2>(t, P, L, bs, nthr, lds, st);
}
template <int RM, int RP>
static void launch_irow_r(const LaunchTab &t, int mode, int slot, int osel, size_t lds, hipStream_t st) {
    const dim3 grid(t.nwg), blk(256);
    switch (mode) {
        case IROW_Z: hipLaunchKernelGGL((k_irow<IROW_Z, RM, RP>), grid, blk, lds, st, KT, slot, osel); break;
        case IROW_FIRST: hipLaunchKernelGGL((k_irow<IROW_FIRST, RM, RP>), grid, blk, lds, st, KT, slot, osel); break;
        default: hipLaunchKernelGGL((k_irow<IROW_UPDATE, RM, RP>), grid, blk, lds, st, KT, slot, osel); break;
    }
}
void launch_irow(const LaunchTab &t, int mode, int rm, int rp, int slot, int osel, size_t lds, hipStream_t st) {
    if (rm > 4) launch_irow_r<7, 0>(t, mode, slot, osel, lds, st);
    else if (rp == 3) launch_irow_r<3, 3>(t, mode, slot, osel, lds, st);
    else if (rp == 2) launch_irow_r<3, 2>(t, mode, slot, osel, lds, st);
    else launch_irow_r<3, 0>(t, mode, slot, osel, lds, st);
}
template <typename RT>
static void launch_wfs_lhs_tt_rt(const LaunchTab &tg, const LaunchTab &tt, int pos, int xsel, hipStream_t st) {
    const dim3 blk(256);
    if (pos == 1) hipLaunchKernelGGL((k_wfsg<1, RT>), dim3(tg.nwg), blk, 0, st, tg.jobs, tg.wgi, xsel);
    else hipLaunchKernelGGL((k_wfsg<2, RT>), dim3(tg.nwg), blk, 0, st, tg.jobs, tg.wgi, xsel);
    hipLaunchKernelGGL((k_ttl<RT>), dim3(tt.nwg), blk, 0, st, tt.jobs, tt.wgi);
    if (pos == 1) hipLaunchKernelGGL((k_wfs<1, true, RT, true>), dim3(tg.nwg), blk, 0, st, tg.jobs, tg.wgi, xsel);
    else hipLaunchKernelGGL((k_wfs<2, true, RT, true>), dim3(tg.nwg), blk, 0, st, tg.jobs, tg.wgi, xsel);
}
void launch_wfs_lhs_tt(const LaunchTab &tg, const LaunchTab &tt, int pos, int rt64, int xsel, hipStream_t st) {
    if (rt64) launch_wfs_lhs_tt_rt<double>(tg, tt, pos, xsel, st);
    else launch_wfs_lhs_tt_rt<float>(tg, tt, pos, xsel, st);
}
template <typename RT>
static void launch_cgv_rt(const LaunchTab &t, int mode, int iter, hipStream_t st) {
    const dim3 grid(t.nwg), blk(256);
    switch (mode) {
        case CGV_INIT_COLD: hipLaunchKernelGGL((k_cgv<CGV_INIT_COLD, RT>), grid, blk, 0, st, KT, iter); break;
        case CGV_INIT_WARM: hipLaunchKernelGGL((k_cgv<CGV_INIT_WARM, RT>), grid, blk, 0, st, KT, iter); break;
        case CGV_UPDATE: hipLaunchKernelGGL((k_cgv<CGV_UPDATE, RT>), grid, blk, 0, st, KT, iter); break;
        default: hipLaunchKernelGGL((k_cgv<CGV_LAST, RT>), grid, blk, 0, st, KT, iter); break;
    }
}
void launch_cgv(const LaunchTab &t, int mode, int rt64, int iter, hipStream_t st) {
    if (rt64) launch_cgv_rt<double>(t, mode, iter, st);
    else launch_cgv_rt<float>(t, mode, iter, st);
}
void launch_cgp(const LaunchTab &t, int mode, int rt64, int slot, hipStream_t st) {
    const dim3 grid(t.nwg), blk(256);
    if (rt64) {
        if (mode == CGP_FIRST) hipLaunchKernelGGL((k_cgp<CGP_FIRST, double>), grid, blk, 0, st, KT, slot);
        else hipLaunchKernelGGL((k_cgp<CGP_UPDATE, double>), grid, blk, 0, st, KT, slot);
    } else {
        if (mode == CGP_FIRST) hipLaunchKernelGGL((k_cgp<CGP_FIRST, float>), grid, blk, 0, st, KT, slot);
        else hipLaunchKernelGGL((k_cgp<CGP_UPDATE, float>), grid, blk, 0, st, KT, slot);
    }
}
void launch_final(const LaunchTab &t, hipStream_t st) {
    hipLaunchKernelGGL(k_final, dim3(t.nwg), dim3(256), 0, st, KT);
}
void launch_zero_stop(const LaunchTab &t, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_stop, dim3(t.nwg), dim3(256), 0, st, KT);
}
void launch_stage_in(const Job *jobs, int njobs, const StageOff *off, const float *src, hipStream_t st) {
    if (njobs > 0) hipLaunchKernelGGL(k_stage_in, dim3((unsigned)(njobs * SB)), dim3(256), 0, st, jobs, off, src);
}
void launch_stage_out(const Job *jobs, int njobs, const StageOff *off, float *dx, double *dres, hipStream_t st) {
    if (njobs > 0)
        hipLaunchKernelGGL(k_stage_out, dim3((unsigned)(njobs * SB)), dim3(256), 0, st, jobs, off, dx, dres);
}
void launch_copy(const float *a, float *b, long n, hipStream_t st) {
    const long nb = (n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)nb), dim3(256), 0, st, a, b, n);
}
void launch_unit(float *v, long len, const int *idx, int n, hipStream_t st) {
    const long nb = (len + 255) / 256 < 256 ? (len + 255) / 256 : 256;
    hipLaunchKernelGGL(k_unit, dim3((unsigned)nb, (unsigned)n), dim3(256), 0, st, v, len, idx);
}

}  // namespace mtk
