// Context, error reporting and small utility kernels of libwbmi.
#include <cstdarg>

#include "wbmi_common.h"

int wbmi_fail(wbmi_ctx* ctx, int code, const char* fmt, ...) {
    if (ctx) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

__global__ void k_zero(uint4* p, int64_t n16, unsigned char* tail, int64_t ntail) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t j = i; j < n16; j += stride) p[j] = make_uint4(0, 0, 0, 0);
    if (i < ntail) tail[i] = 0;
}

int wbmi_zero(wbmi_ctx* ctx, void* ptr, size_t bytes) {
    if (bytes == 0) return WBMI_OK;
    if (((uintptr_t)ptr) % 16 != 0) return wbmi_fail(ctx, WBMI_ERR_ARG, "wbmi_zero: pointer not 16-byte aligned");
    int64_t n16 = bytes / 16;
    int64_t ntail = bytes % 16;
    int64_t blocks = ceil_div(n16 > 0 ? n16 : 1, 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, (uint4*)ptr, n16,
                       (unsigned char*)ptr + n16 * 16, ntail);
    WBMI_LAUNCH_CHECK(ctx);
    return WBMI_OK;
}

extern "C" {

int wbmi_version(void) { return 1; }

int wbmi_ctx_create(int device, void* stream, wbmi_ctx** out) {
    if (!out) return WBMI_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return WBMI_ERR_HIP;
    if (device < 0 || device >= ndev) return WBMI_ERR_ARG;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return WBMI_ERR_HIP;
    if (std::string(prop.gcnArchName).find("gfx950") == std::string::npos) return WBMI_ERR_UNSUPPORTED;
    wbmi_ctx* c = new wbmi_ctx();
    c->device = device;
    c->stream = (hipStream_t)stream;
    c->eigh_scratch = nullptr;
    // the scratch array's address on this device, looked up with the device current
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess) {
        delete c;
        return WBMI_ERR_HIP;
    }
    const int rc = wbmi_eigh_scratch_init(&c->eigh_scratch);
    if (hipSetDevice(prev) != hipSuccess || rc != 0) {
        delete c;
        return WBMI_ERR_HIP;
    }
    *out = c;
    return WBMI_OK;
}

int wbmi_ctx_destroy(wbmi_ctx* ctx) {
    delete ctx;
    return WBMI_OK;
}

int wbmi_ctx_set_stream(wbmi_ctx* ctx, void* stream) {
    if (!ctx) return WBMI_ERR_ARG;
    ctx->stream = (hipStream_t)stream;
    return WBMI_OK;
}

int wbmi_sync(wbmi_ctx* ctx) {
    if (!ctx) return WBMI_ERR_ARG;
    WBMI_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    WBMI_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return WBMI_OK;
}

const char* wbmi_last_error(const wbmi_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

}  // extern "C"
