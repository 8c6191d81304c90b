// capi_common.cpp -- thread-local error state and HIP helpers of the C ABI.
#include "capi_common.hpp"

#include <string>

namespace tfhe {

static thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

int hipcheck(hipError_t e, const char* what) {
    if (e == hipSuccess) return TFHE_OK;
    return fail(TFHE_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

int check_launch(const char* what) { return hipcheck(hipGetLastError(), what); }

int dmalloc(void** p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) return TFHE_OK;
    return hipcheck(hipMalloc(p, bytes), "hipMalloc");
}

}  // namespace tfhe

extern "C" const char* tfhe_last_error(void) { return tfhe::g_last_error.c_str(); }
