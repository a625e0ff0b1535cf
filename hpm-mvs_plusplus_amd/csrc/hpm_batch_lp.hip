// hpm_batch_lp.hip -- the fused multi-view schedule's FP16 and BF16 engine
// instantiations (compiled beside hpm_batch.hip).
#include "hpm_batch_impl.h"

template int hpm_batch_detail::run_engine<FastFor<hpm::PairF16>, true>(hpm_batch*, int, int);
template int hpm_batch_detail::run_engine<FastFor<hpm::PairU8>, true>(hpm_batch*, int, int);
template int hpm_batch_detail::run_engine<FastFor<hpm::PairBF16>, true>(hpm_batch*, int, int);
