// mkde_rates.hip -- instruction-rate probe for the roofline of the KDE kernels
// on gfx950 (no reference counterpart: measurement only).
//
// Every Gaussian evaluation of the path costs FP32 VALU work and one
// exponential (v_exp_f32); the density raster also runs FP32 MFMA.  This
// program measures, on the device it runs on:
//   exp   lane v_exp_f32 per second (8 independent chains per lane)
//   fma   lane v_fmac_f32 per second (8 independent chains per lane)
//   mix   exp and fma chains interleaved 1:R: if the two issue side by side
//         the time is max(t_exp, t_fma), if they share the issue port it is
//         t_exp + t_fma; the reported co-issue fraction is
//         (t_exp + t_fma - t_mix) / min(t_exp, t_fma)
//   mfma  v_mfma_f32_32x32x2_f32 FLOP per second (4 independent accumulators)
// and prints one JSON line.  Each kernel's wave count fills the device several
// times over (8 waves per SIMD).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>

#define CHAINS 8

#define HIPCHK(x)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (x);                                                             \
        if (e_ != hipSuccess) {                                                          \
            fprintf(stderr, "%s failed: %s\n", #x, hipGetErrorString(e_));               \
            exit(1);                                                                     \
        }                                                                                \
    } while (0)

// v <- 2^-v: converges to a fixed point, one v_exp_f32 per step (the negation
// is a source modifier).
__global__ __launch_bounds__(256) void k_exp(float *out, int iters)
{
    float v[CHAINS];
#pragma unroll
    for (int j = 0; j < CHAINS; ++j) v[j] = 1e-3f * (float)threadIdx.x + 0.1f * (float)j;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < CHAINS; ++j) v[j] = __builtin_amdgcn_exp2f(-v[j]);
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < CHAINS; ++j) s += v[j];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// v <- v * a + b: one v_fmac_f32 per step.
__global__ __launch_bounds__(256) void k_fma(float *out, int iters, float a, float b)
{
    float v[CHAINS];
#pragma unroll
    for (int j = 0; j < CHAINS; ++j) v[j] = 1e-3f * (float)threadIdx.x + 0.1f * (float)j;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < CHAINS; ++j) v[j] = __builtin_fmaf(v[j], a, b);
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < CHAINS; ++j) s += v[j];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// CHAINS exp chains and CHAINS fma chains, R fma steps per exp step.
template <int R>
__global__ __launch_bounds__(256) void k_mix(float *out, int iters, float a, float b)
{
    float v[CHAINS], u[CHAINS];
#pragma unroll
    for (int j = 0; j < CHAINS; ++j) {
        v[j] = 1e-3f * (float)threadIdx.x + 0.1f * (float)j;
        u[j] = v[j] + 0.5f;
    }
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < CHAINS; ++j) {
            v[j] = __builtin_amdgcn_exp2f(-v[j]);
#pragma unroll
            for (int r = 0; r < R; ++r) u[j] = __builtin_fmaf(u[j], a, b);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < CHAINS; ++j) s += v[j] + u[j];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

typedef float f32x16 __attribute__((ext_vector_type(16)));

// 4 independent 32x32 accumulators, one v_mfma_f32_32x32x2_f32 each per step.
__global__ __launch_bounds__(256) void k_mfma(float *out, int iters, float a, float b)
{
    f32x16 c0 = {}, c1 = {}, c2 = {}, c3 = {};
    const float x = a + 1e-6f * (float)threadIdx.x, y = b - 1e-6f * (float)threadIdx.x;
    for (int i = 0; i < iters; ++i) {
        c0 = __builtin_amdgcn_mfma_f32_32x32x2f32(x, y, c0, 0, 0, 0);
        c1 = __builtin_amdgcn_mfma_f32_32x32x2f32(y, x, c1, 0, 0, 0);
        c2 = __builtin_amdgcn_mfma_f32_32x32x2f32(x, x, c2, 0, 0, 0);
        c3 = __builtin_amdgcn_mfma_f32_32x32x2f32(y, y, c3, 0, 0, 0);
    }
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < 16; ++j) s += c0[j] + c1[j] + c2[j] + c3[j];
    out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

struct Timer {
    hipEvent_t a, b;
    Timer()
    {
        HIPCHK(hipEventCreate(&a));
        HIPCHK(hipEventCreate(&b));
    }
    ~Timer()
    {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

template <class F>
static double time_ms(F launch, int reps)
{
    Timer t;
    launch();   // warm-up
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    float best = 1e30f;
    for (int r = 0; r < reps; ++r) {
        HIPCHK(hipEventRecord(t.a, 0));
        launch();
        HIPCHK(hipEventRecord(t.b, 0));
        HIPCHK(hipEventSynchronize(t.b));
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, t.a, t.b));
        if (ms < best) best = ms;
    }
    return best;
}

int main(int argc, char **argv)
{
    int dev = argc > 1 ? atoi(argv[1]) : 0;
    HIPCHK(hipSetDevice(dev));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, dev));
    const int cus = prop.multiProcessorCount;
    const int blocks = cus * 8;   // 8 workgroups of 4 waves per CU: 8 waves per SIMD
    const int threads = 256;
    const size_t nthr = (size_t)blocks * threads;
    float *out = nullptr;
    HIPCHK(hipMalloc((void **)&out, nthr * sizeof(float)));
    const int iters = 4096, reps = 5;
    const double lane_steps = (double)nthr * iters * CHAINS;

    const double t_exp = time_ms([&] { hipLaunchKernelGGL(k_exp, dim3(blocks), dim3(threads), 0, 0, out, iters); },
                                 reps);
    const double t_fma = time_ms(
        [&] { hipLaunchKernelGGL(k_fma, dim3(blocks), dim3(threads), 0, 0, out, iters, 0.999f, 0.001f); }, reps);
    const double t_mix1 = time_ms(
        [&] { hipLaunchKernelGGL(k_mix<1>, dim3(blocks), dim3(threads), 0, 0, out, iters, 0.999f, 0.001f); }, reps);
    const double t_mix4 = time_ms(
        [&] { hipLaunchKernelGGL(k_mix<4>, dim3(blocks), dim3(threads), 0, 0, out, iters, 0.999f, 0.001f); }, reps);
    const double t_fma4 = t_fma * 4.0;
    const int mfma_iters = 2048;
    const double t_mfma = time_ms(
        [&] { hipLaunchKernelGGL(k_mfma, dim3(blocks), dim3(threads), 0, 0, out, mfma_iters, 0.5f, 0.25f); }, reps);
    const double mfma_flop = (double)blocks * (threads / 64) * mfma_iters * 4 * (32.0 * 32.0 * 2.0 * 2.0);

    const double exp_rate = lane_steps / (t_exp * 1e-3);
    const double fma_rate = lane_steps / (t_fma * 1e-3);
    auto coissue = [](double te, double tf, double tm) {
        const double m = te < tf ? te : tf;
        return m > 0.0 ? (te + tf - tm) / m : 0.0;
    };
    printf("{\"device\": \"%s\", \"cus\": %d, \"clock_mhz\": %d, "
           "\"exp_per_s\": %.6e, \"fma_per_s\": %.6e, \"fp32_valu_tflops\": %.3f, "
           "\"exp_to_fma_ratio\": %.4f, "
           "\"mix_1to1_ms\": %.4f, \"exp_ms\": %.4f, \"fma_ms\": %.4f, \"coissue_1to1\": %.3f, "
           "\"mix_1to4_ms\": %.4f, \"coissue_1to4\": %.3f, "
           "\"mfma_f32_32x32x2_tflops\": %.3f}\n",
           prop.gcnArchName, cus, prop.clockRate / 1000, exp_rate, fma_rate, 2.0 * fma_rate / 1e12,
           exp_rate / fma_rate, t_mix1, t_exp, t_fma, coissue(t_exp, t_fma, t_mix1), t_mix4,
           coissue(t_exp, t_fma4, t_mix4), mfma_flop / (t_mfma * 1e-3) / 1e12);
    HIPCHK(hipFree(out));
    return 0;
}
