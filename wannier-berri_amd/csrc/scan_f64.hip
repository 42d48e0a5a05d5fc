// Fermi-bin scan kernels for complex128 channels (instantiations of scan_kernels.h).
#include "scan_kernels.h"

int scan_dispatch_f64(wbmi_ctx* ctx, int nk, const ScanArgs& a, const double* E, const void* X, double* hist) {
    return dispatch_scan<double, 0>(ctx, nk, a, E, X, hist);
}
