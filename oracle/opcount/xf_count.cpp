/*
 * Counting glue (test infrastructure): counters, and the deterministic
 * exp/log of ora_math.c (built as plain C doubles under *_plain names)
 * re-exported for the counted sources as one counted operation each.
 */
#include "xf_count.hpp"
#undef double

extern "C" {
unsigned long long xf_cnt[8];
double ora_xexp_plain(double);
double ora_xlog_plain(double);

void ora_count_reset(void) { memset(xf_cnt, 0, sizeof xf_cnt); }
void ora_count_read(unsigned long long *out) { memcpy(out, xf_cnt, sizeof xf_cnt); }
xf_cd ora_exp(xf_cd x) { xf_cnt[4]++; return xf_cd(ora_xexp_plain(x.v)); }
xf_cd ora_log(xf_cd x) { xf_cnt[5]++; return xf_cd(ora_xlog_plain(x.v)); }
}

xf_cd ora_xexp(xf_cd x) { xf_cnt[4]++; return xf_cd(ora_xexp_plain(x.v)); }
xf_cd ora_xlog(xf_cd x) { xf_cnt[5]++; return xf_cd(ora_xlog_plain(x.v)); }
