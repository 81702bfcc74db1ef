/* Correctly rounded natural logarithm for the oracle (TEST INFRASTRUCTURE).
 *
 * The reference takes log() from the host libm in the anomalous-resistivity
 * mask of its CT step (src/prj_timeint.c:230-263).  A libm log is not
 * required to round correctly (glibc's does not on a small fraction of
 * inputs), and the GPU path rounds it correctly.  fxo_log_cr evaluates log in
 * binary128 (libquadmath, ~2^-112 relative error) and rounds once to double,
 * which is the correctly rounded value unless log(x) lies within ~2^-110
 * relative of a rounding boundary.  Independent of the GPU's double-double
 * method, so the two check each other. */
#include <quadmath.h>

double fxo_log_cr(double x)
{
    return (double)logq((__float128)x);
}
