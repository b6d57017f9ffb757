// Bidiagonalization of large single matrices: placeholder that routes to the
// one-workgroup kernel until the grid-cooperative panel kernel is in place.
#include "paco_kernels.hpp"

namespace paco {

size_t bidiag_grid_ws_bytes(int M, int N)
{
    (void)M;
    (void)N;
    return 0;
}

hipError_t launch_bidiag_grid(const ProbDesc &P, void *ws, hipStream_t s)
{
    (void)ws;
    ProbList pl;
    pl.ptr = nullptr;
    pl.single = P;
    pl.count = 1;
    return launch_bidiag_wg(pl, nullptr, 1, bidiag_wg_lds_bytes(P.M, P.N), s);
}

} // namespace paco
