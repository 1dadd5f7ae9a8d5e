/*
 * backward.cpp - xm_contract_backward (no libxm counterpart).
 */
#include "contract_plan.h"

XM_API void
xm_contract_backward(xm_scalar_t, const xm_tensor_t *, const xm_tensor_t *,
    const xm_tensor_t *, xm_tensor_t *, xm_tensor_t *, const char *,
    const char *, const char *)
{
	XM_FATAL("xm_contract_backward", "not available");
}
