// Streaming-bandwidth probe: the practical HBM ceiling of plain streams on
// one GPU, against which the time-step kernels' measured traffic is read.
// Three streams over 512 MiB arrays, float4 per lane:
//   copy   one read, one write      (2 arrays)
//   r4w1   four reads, one write    (the time step's read/write mix)
//   read   reads only               (one float per lane written at the end)
// each launched as a wide grid (one float4 per lane) and as grid-stride grids
// of 1024..32768 blocks; every access is inside its array (i < n).  Prints
// one JSON line per stream and grid, then the best of each stream.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

#define CK(x)                                                                                  \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) {                                                                \
            fprintf(stderr, "hbm_probe: %s at line %d\n", hipGetErrorString(e_), __LINE__);    \
            exit(1);                                                                           \
        }                                                                                      \
    } while (0)

__global__ void k_copy(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ void k_r4w1(const float4 *__restrict__ a, const float4 *__restrict__ b, const float4 *__restrict__ c,
                       const float4 *__restrict__ d, float4 *__restrict__ o, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float4 x = a[i], y = b[i], z = c[i], w = d[i];
        o[i] = make_float4(x.x + y.x + z.x + w.x, x.y + y.y + z.y + w.y, x.z + y.z + z.z + w.z, x.w + y.w + z.w + w.w);
    }
}

__global__ void k_read(const float4 *__restrict__ a, float *__restrict__ o, size_t n)
{
    float s = 0.0f;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float4 x = a[i];
        s += (x.x + x.y) + (x.z + x.w);
    }
    o[blockIdx.x * (size_t)blockDim.x + threadIdx.x] = s;
}

int main()
{
    const size_t bytes = 512ull << 20, n = bytes / sizeof(float4);
    const int nt = 256;
    const int wide = (int)(n / nt);   // one float4 per lane
    float4 *A[5];
    for (int k = 0; k < 5; ++k) {
        CK(hipMalloc(&A[k], bytes));
        CK(hipMemset(A[k], 0, bytes));
    }
    float *red;   // one float per lane of the widest grid
    CK(hipMalloc(&red, sizeof(float) * (size_t)wide * nt));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int grids[] = {1024, 4096, 32768, wide};
    const char *names[] = {"copy", "r4w1", "read"};
    const double moved[] = {2.0, 5.0, 1.0};   // arrays touched per element
    double best[3] = {0.0, 0.0, 0.0};
    const int reps = 20;
    for (int kind = 0; kind < 3; ++kind) {
        for (int g : grids) {
            auto go = [&]() {
                if (kind == 0)
                    k_copy<<<g, nt>>>(A[0], A[1], n);
                else if (kind == 1)
                    k_r4w1<<<g, nt>>>(A[0], A[1], A[2], A[3], A[4], n);
                else
                    k_read<<<g, nt>>>(A[0], red, n);
            };
            for (int r = 0; r < 3; ++r) go();
            CK(hipGetLastError());
            CK(hipEventRecord(e0));
            for (int r = 0; r < reps; ++r) go();
            CK(hipEventRecord(e1));
            CK(hipEventSynchronize(e1));
            float ms = 0.0f;
            CK(hipEventElapsedTime(&ms, e0, e1));
            const double tbs = moved[kind] * (double)bytes * reps / (ms * 1e-3) / 1e12;
            if (tbs > best[kind]) best[kind] = tbs;
            printf("{\"stream\": \"%s\", \"blocks\": %d, \"bytes_per_launch\": %.0f, \"TB_per_s\": %.3f}\n", names[kind], g,
                   moved[kind] * (double)bytes, tbs);
            fflush(stdout);
        }
    }
    printf("{\"best_TB_per_s\": {\"copy\": %.3f, \"r4w1\": %.3f, \"read\": %.3f}}\n", best[0], best[1], best[2]);
    for (int k = 0; k < 5; ++k) CK(hipFree(A[k]));
    CK(hipFree(red));
    return 0;
}
