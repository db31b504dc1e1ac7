// hbm_copy: calibration for rocprofv3's FETCH_SIZE / WRITE_SIZE counters and
// a measured streaming ceiling.
//
// copy_f32 copies N floats with 4-byte accesses (reads exactly N*4 bytes,
// writes N*4 bytes per launch): comparing the counters' per-dispatch values
// with those sizes gives the factors that turn counter readings into HBM
// bytes.  copy_v4 copies the same arrays with 16-byte accesses, four per
// thread per grid-stride step, over arrays from 256 MiB to 4 GiB (beyond
// the 256 MiB of last-level cache), on a grid of 8 blocks per CU: the best
// of these is the practical HBM ceiling the roofline fractions are read
// against beside the nominal 8 TB/s.
#include <hip/hip_runtime.h>

#include <cstdio>

__global__ void copy_f32(const float *__restrict__ a, float *__restrict__ b, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) b[i] = a[i];
}

// n4: float4 elements, a multiple of 4 * blockDim.x * gridDim.x
__global__ __launch_bounds__(256) void copy_v4(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n4)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += 4 * stride) {
        const float4 x0 = a[i], x1 = a[i + stride], x2 = a[i + 2 * stride], x3 = a[i + 3 * stride];
        b[i] = x0;
        b[i + stride] = x1;
        b[i + 2 * stride] = x2;
        b[i + 3 * stride] = x3;
    }
}

static float time_launches(int reps, void (*launch)(void *), void *arg)
{
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return -1.f;
    for (int it = 0; it < 3; it++) launch(arg);
    (void)hipEventRecord(e0);
    for (int it = 0; it < reps; it++) launch(arg);
    (void)hipEventRecord(e1);
    float ms = 0.f;
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) ms = -1.f;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms;
}

struct CopyArgs {
    float *a, *b;
    size_t n;
    int grid;
};

static void launch_f32(void *p)
{
    const CopyArgs *c = (const CopyArgs *)p;
    hipLaunchKernelGGL(copy_f32, dim3(8192), dim3(256), 0, 0, c->a, c->b, c->n);
}

static void launch_v4(void *p)
{
    const CopyArgs *c = (const CopyArgs *)p;
    hipLaunchKernelGGL(copy_v4, dim3(c->grid), dim3(256), 0, 0, (const float4 *)c->a, (float4 *)c->b, c->n / 4);
}

int main()
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        return 1;
    const size_t nmax = 1024ull << 20;  // floats: 4 GiB per array
    float *a, *b;
    if (hipMalloc(&a, nmax * 4) != hipSuccess || hipMalloc(&b, nmax * 4) != hipSuccess) return 1;
    if (hipMemset(a, 0, nmax * 4) != hipSuccess || hipMemset(b, 0, nmax * 4) != hipSuccess) return 1;
    // the counter calibration: 256 MiB per array, 4-byte accesses
    CopyArgs c{a, b, 64ull << 20, 8 * cus};
    float ms = time_launches(10, launch_f32, &c);
    printf("copy_f32 %zu MiB x2 per launch: %.1f GB/s\n", (c.n * 4) >> 20, 10 * 2.0 * c.n * 4 / (ms * 1e-3) / 1e9);
    // the streaming ceiling: 16-byte accesses, 256 MiB .. 4 GiB per array
    double best = 0.0;
    for (size_t n = 64ull << 20; n <= nmax; n *= 2) {
        c.n = n;
        c.grid = 8 * cus;
        const size_t per = 4ull * 256 * c.grid * 4;  // floats per grid-stride step
        c.n = (n / per) * per;
        ms = time_launches(10, launch_v4, &c);
        const double gbs = 10 * 2.0 * c.n * 4 / (ms * 1e-3) / 1e9;
        if (gbs > best) best = gbs;
        printf("copy_v4 %zu MiB x2 per launch: %.1f GB/s\n", (c.n * 4) >> 20, gbs);
    }
    printf("best copy_v4: %.1f GB/s\n", best);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}
