// Probe: the cost of a grid-wide barrier on this GPU (cooperative groups and a
// hand-written counter barrier) against kernel launches and graph kernel
// nodes, for weighing a persistent one-launch PCG solve against a captured
// graph of five launches per iteration.  Build and run:
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         -o grid_barrier grid_barrier.hip && ./grid_barrier
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>
#include <cstdio>
#include <vector>
namespace cg = cooperative_groups;

__global__ void k_bar(int iters, float *buf, int touch) {
    cg::grid_group g = cg::this_grid();
    float acc = 0.f;
    for (int i = 0; i < iters; i++) {
        if (touch) {
            // each phase writes one value per workgroup and reads a neighbour's
            if (threadIdx.x == 0) buf[blockIdx.x] = (float)i + acc;
        }
        g.sync();
        if (touch && threadIdx.x == 0) acc += buf[(blockIdx.x + 1) % gridDim.x];
    }
    if (threadIdx.x == 0) buf[gridDim.x + blockIdx.x] = acc;
}

// Hand-written grid barrier: one arrival counter and a generation word; the
// last workgroup to arrive resets the counter and advances the generation.
// Every wait is bounded: after `limit` polls a workgroup sets *err and the
// later barriers of the launch no longer wait.
__device__ __forceinline__ void my_sync(unsigned *count, unsigned *gen, unsigned nb, unsigned &mygen,
                                        unsigned limit, unsigned *err) {
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned g = mygen;
        if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
            const unsigned a = __hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (a == nb - 1) {
                __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(gen, g + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                unsigned n = 0;
                while (__hip_atomic_load(gen, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == g && ++n < limit)
                    __builtin_amdgcn_s_sleep(1);
                if (n >= limit) __hip_atomic_fetch_or(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        mygen = g + 1;
    }
    __syncthreads();
}

__global__ void k_mybar(int iters, float *buf, int touch, unsigned *ctl, unsigned gen0) {
    unsigned mygen = gen0;
    float acc = 0.f;
    for (int i = 0; i < iters; i++) {
        if (touch && threadIdx.x == 0) buf[blockIdx.x] = (float)i + acc;
        my_sync(ctl, ctl + 32, gridDim.x, mygen, 1u << 20, ctl + 64);
        if (touch && threadIdx.x == 0) acc += buf[(blockIdx.x + 1) % gridDim.x];
    }
    if (threadIdx.x == 0) buf[gridDim.x + blockIdx.x] = acc;
}

__global__ void k_empty(float *buf) {
    if (threadIdx.x == 0 && blockIdx.x == 0) buf[0] += 1.f;
}

#define CK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d: %s\n", #x, __LINE__, hipGetErrorString(e)); return 1; } } while (0)

int main() {
    int dev = 0, coop = 0, ncu = 0;
    CK(hipSetDevice(dev));
    CK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
    CK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    int per = 0;
    CK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_bar, 256, 0));
    printf("cooperative=%d CUs=%d max blocks/CU(256 thr)=%d\n", coop, ncu, per);
    float *buf;
    CK(hipMalloc(&buf, sizeof(float) * 65536));
    CK(hipMemset(buf, 0, sizeof(float) * 65536));
    hipEvent_t a, b;
    CK(hipEventCreate(&a));
    CK(hipEventCreate(&b));
    const int iters = 2000;
    for (int touch = 0; touch < 2; touch++) {
        for (int nb : {1, 8, 64, 256, 512, 1024}) {
            if (nb > ncu * per) continue;
            int it = iters;
            void *args[] = {&it, &buf, &touch};
            // warm-up
            CK(hipLaunchCooperativeKernel((void *)k_bar, dim3(nb), dim3(256), args, 0, 0));
            CK(hipDeviceSynchronize());
            CK(hipEventRecord(a, 0));
            CK(hipLaunchCooperativeKernel((void *)k_bar, dim3(nb), dim3(256), args, 0, 0));
            CK(hipEventRecord(b, 0));
            CK(hipEventSynchronize(b));
            float ms = 0;
            CK(hipEventElapsedTime(&ms, a, b));
            printf("grid barrier: blocks=%5d touch=%d  %.3f us per barrier\n", nb, touch, 1e3 * ms / iters);
            fflush(stdout);
        }
    }

    // hand-written barrier; ctl: [0] counter, [32] generation, [64] error
    unsigned *ctl;
    CK(hipMalloc(&ctl, 4096));
    CK(hipMemset(ctl, 0, 4096));
    unsigned gen = 0;
    for (int touch = 0; touch < 2; touch++) {
        for (int nb : {1, 8, 64, 256, 512, 1024}) {
            if (nb > ncu * per) continue;
            int it = iters;
            for (int rep = 0; rep < 2; rep++) {
                void *args[] = {&it, &buf, &touch, &ctl, &gen};
                CK(hipEventRecord(a, 0));
                CK(hipLaunchCooperativeKernel((void *)k_mybar, dim3(nb), dim3(256), args, 0, 0));
                CK(hipEventRecord(b, 0));
                CK(hipEventSynchronize(b));
                gen += iters;
                float ms = 0;
                CK(hipEventElapsedTime(&ms, a, b));
                unsigned herr = 0;
                CK(hipMemcpy(&herr, ctl + 64, 4, hipMemcpyDeviceToHost));
                if (rep == 1)
                    printf("atomic barrier: blocks=%5d touch=%d  %.3f us per barrier  err=%u\n", nb, touch,
                           1e3 * ms / iters, herr);
                fflush(stdout);
                if (herr) { printf("barrier wait limit reached; stopping\n"); return 1; }
            }
        }
    }
    CK(hipFree(ctl));
    // launch cost for comparison: back-to-back empty kernels on one stream,
    // and the same captured in a graph
    const int nl = 2000;
    CK(hipEventRecord(a, 0));
    for (int i = 0; i < nl; i++) hipLaunchKernelGGL(k_empty, dim3(256), dim3(256), 0, 0, buf);
    CK(hipEventRecord(b, 0));
    CK(hipEventSynchronize(b));
    float ms = 0;
    CK(hipEventElapsedTime(&ms, a, b));
    printf("back-to-back launches (256 blocks): %.3f us per launch\n", 1e3 * ms / nl);
    hipStream_t s;
    CK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    hipGraph_t gr;
    hipGraphExec_t ge;
    CK(hipStreamBeginCapture(s, hipStreamCaptureModeGlobal));
    for (int i = 0; i < 200; i++) hipLaunchKernelGGL(k_empty, dim3(256), dim3(256), 0, s, buf);
    CK(hipStreamEndCapture(s, &gr));
    CK(hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0));
    CK(hipGraphLaunch(ge, s));
    CK(hipStreamSynchronize(s));
    CK(hipEventRecord(a, s));
    for (int r = 0; r < 10; r++) CK(hipGraphLaunch(ge, s));
    CK(hipEventRecord(b, s));
    CK(hipEventSynchronize(b));
    CK(hipEventElapsedTime(&ms, a, b));
    printf("graph of 200 launches (256 blocks): %.3f us per kernel node\n", 1e3 * ms / 2000);
    CK(hipFree(buf));
    printf("done\n");
    return 0;
}
