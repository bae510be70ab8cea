// C-callable entry into the reference's own LuksanPLIS (potential/LuksanPLIS.h:35-36), so that tests can
// drive it through ctypes.  Test infrastructure only: built into oracle/_ref, never into the product.
#include "LuksanPLIS.h"

extern "C" int ref_luksan_plis(int n, LuksanObjective f, void *f_data, double *x, double *minf, double stopval,
                               int maxeval, int mf, int *evaluations, int fast)
{
	return LuksanPLIS(n, f, f_data, x, minf, stopval, maxeval, mf, evaluations, fast);
}
