// tetrahedron-weighted scan kernels for complex64 channels (instantiations of scan_kernels.h).
#include "scan_kernels.h"

int scan_dispatch_tetra_f32(wbmi_ctx* ctx, int nk, const ScanArgs& a, const double* E, const void* X, double* hist) {
    return dispatch_scan<float, 1>(ctx, nk, a, E, X, hist);
}
