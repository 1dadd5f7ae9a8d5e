/*
 * split.hip - split-precision (BF16 slices) FP64 emulation for
 * xm_contract.  This stage is not built yet: requesting SPLIT_BF16
 * fails loudly instead of silently computing in another precision.
 */
#include "contract_plan.h"

void
split_prepare(contract_plan *pl, const xm_tensor_t *, const xm_tensor_t *,
    const xm_tensor_t *, const masks_t &)
{
	(void)pl;
	XM_FATAL("xm_contract", "split-precision contraction not available");
}

void split_slices(contract_plan *, hipStream_t) {}
void split_gemm(contract_plan *, hipStream_t) {}
void split_release(contract_plan *) {}
