/* hd/default_params.h -- replaces Pipeline/hd/default_params.h:16 */
#pragma once
#include "hd/params.h"
#ifdef __cplusplus
extern "C" {
#endif
/* replaces hd_set_default_params (Pipeline/default_params.C:10-66) */
void hd_set_default_params(struct hd_params* params);
#ifdef __cplusplus
}
#endif
