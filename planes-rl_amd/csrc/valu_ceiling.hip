// valu_ceiling -- measures the FP32 vector-ALU issue ceiling of the device the flight kernel
// runs on: wave-instructions per cycle per SIMD for chains of v_fma_f32 (and v_mul/v_add,
// the v_rsq_f32 / v_sqrt_f32 the flight step issues twice per panel, and a 3:1 FMA:rsq mix),
// dependent and independent, at 1..8 waves per SIMD, with the shader clock read from
// s_memtime / s_memrealtime (100 MHz).  The flight kernel is built without packed FP32, so
// its ceiling is (this rate) x 64 lanes x 2 FLOP x SIMDs x clock, not the packed-FP32 peak.
#include <hip/hip_runtime.h>

#include <cstdio>

template <int C, int OP>
__global__ __launch_bounds__(64) void k_chain(float* out, long long* clk, float a, float b, int iters) {
    float x[C];
#pragma unroll
    for (int c = 0; c < C; c++) x[c] = 1.0f + threadIdx.x * 1e-3f + c;
    const long long t0 = __builtin_amdgcn_s_memtime();
    const long long r0 = __builtin_amdgcn_s_memrealtime();
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int u = 0; u < 16; u++) {
#pragma unroll
            for (int c = 0; c < C; c++) {
                if (OP == 0) x[c] = __builtin_fmaf(x[c], a, b);
                else if (OP == 1) x[c] = x[c] * a;
                else if (OP == 2) x[c] = x[c] + b;
                else if (OP == 3) x[c] = __builtin_amdgcn_rsqf(x[c]);  // positive, tends to 1
                else if (OP == 4) x[c] = __builtin_amdgcn_sqrtf(x[c]);
                else x[c] = c % 4 == 0 ? __builtin_amdgcn_rsqf(x[c]) : __builtin_fmaf(x[c], a, b);  // 1 in 4
            }
        }
    }
    const long long t1 = __builtin_amdgcn_s_memtime();
    const long long r1 = __builtin_amdgcn_s_memrealtime();
    float s = 0;
#pragma unroll
    for (int c = 0; c < C; c++) s += x[c];
    out[blockIdx.x * 64 + threadIdx.x] = s;  // grid <= 32 blocks per CU: within the buffer
    if (threadIdx.x == 0 && blockIdx.x == 0) { clk[0] = t1 - t0; clk[1] = r1 - r0; }
}

static bool ok(hipError_t e) {
    if (e != hipSuccess) fprintf(stderr, "HIP error: %s\n", hipGetErrorString(e));
    return e == hipSuccess;
}

template <int C, int OP>
static bool run(const char* name, int wps, float* d, long long* dc, int ncu) {
    const int blocks = ncu * 4 * wps;
    const int iters = 40000 / wps;
    hipEvent_t e0, e1;
    if (!ok(hipEventCreate(&e0)) || !ok(hipEventCreate(&e1))) return false;
    if (!ok(hipEventRecord(e0))) return false;
    hipLaunchKernelGGL((k_chain<C, OP>), dim3(blocks), dim3(64), 0, 0, d, dc, 0.999f, 1e-3f, iters);
    if (!ok(hipGetLastError()) || !ok(hipEventRecord(e1)) || !ok(hipEventSynchronize(e1))) return false;
    float ms = 0;
    long long c[2];
    if (!ok(hipEventElapsedTime(&ms, e0, e1)) || !ok(hipMemcpy(c, dc, sizeof c, hipMemcpyDeviceToHost))) return false;
    const double ghz = (double)c[0] / ((double)c[1] / 100e6) / 1e9;
    const double instr = (double)blocks * iters * 16 * C;
    const double per_cyc = instr / (ncu * 4.0) / (ms * 1e-3 * ghz * 1e9);
    const double tflops = per_cyc * 64 * (OP == 0 ? 2 : OP == 5 ? 1.75 : 1) * ncu * 4.0 * ghz * 1e9 / 1e12;
    printf("%-4s chains=%d waves/SIMD=%d: %7.2f ms, clock %.2f GHz, %.3f wave-instr/cycle/SIMD, %.1f TFLOP/s\n", name,
           C, wps, ms, ghz, per_cyc, tflops);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return true;
}

int main() {
    int ncu = 0;
    if (!ok(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, 0))) return 1;
    float* d = nullptr;
    long long* dc = nullptr;
    if (!ok(hipMalloc(&d, sizeof(float) * 64 * ncu * 4 * 8)) || !ok(hipMalloc(&dc, 16))) return 1;
    bool good = true;
    for (int k = 0; k < 4 && good; k++) good = run<8, 0>("warm", 2, d, dc, ncu);  // clock ramp
    for (int w : {1, 2, 4, 8}) {
        if (!good) break;
        good = run<1, 0>("fma", w, d, dc, ncu) && run<8, 0>("fma", w, d, dc, ncu) &&
               run<8, 1>("mul", w, d, dc, ncu) && run<8, 2>("add", w, d, dc, ncu) &&
               run<8, 3>("rsq", w, d, dc, ncu) && run<8, 4>("sqrt", w, d, dc, ncu) &&
               run<8, 5>("mix", w, d, dc, ncu);
    }
    (void)hipFree(d);
    (void)hipFree(dc);
    return good ? 0 : 1;
}
