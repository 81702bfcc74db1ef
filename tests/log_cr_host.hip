// Host build of the library's correctly rounded log (fornaxamr_amd/csrc/fx_log.hpp),
// compiled by tests/test_log_cr.py to check it on the CPU against exact
// decimal evaluation.  The device build runs the same IEEE operations.
#include "../fornaxamr_amd/csrc/fx_log.hpp"

extern "C" void fx_test_log_cr(int n, const double* x, double* y)
{
    for (int i = 0; i < n; ++i) y[i] = fx::log_cr(x[i]);
}
