// Instantiations and launcher of the elastic-net kernel: 1/3/5/10 column slots per lane
// (n <= 64/192/320/640) x 4/6/8 signal rows per lane (m <= 256/384/512).
#include "lasso_wave.hpp"

namespace lsd {

namespace {

template <int NCPL>
int launch_lasso_rows(const char* name, int rpl, const LassoArgs& a, int waves,
                      hipStream_t stream) {
    switch (rpl) {
    case 4: return launch_fill(name, lasso_kernel<NCPL, 4>, a, waves, 0, stream);
    case 6: return launch_fill(name, lasso_kernel<NCPL, 6>, a, waves, 0, stream);
    case 8: return launch_fill(name, lasso_kernel<NCPL, 8>, a, waves, 0, stream);
    default: return fail("no elastic-net kernel for " + std::to_string(rpl) + " rows per lane");
    }
}

} // namespace

int launch_lasso(const char* name, int ncpl, int rpl, const LassoArgs& a, int waves,
                 hipStream_t stream) {
    switch (ncpl) {
    case 1: return launch_lasso_rows<1>(name, rpl, a, waves, stream);
    case 3: return launch_lasso_rows<3>(name, rpl, a, waves, stream);
    case 5: return launch_lasso_rows<5>(name, rpl, a, waves, stream);
    case 10: return launch_lasso_rows<10>(name, rpl, a, waves, stream);
    default: return fail("no elastic-net kernel for " + std::to_string(ncpl) + " column slots");
    }
}

} // namespace lsd
