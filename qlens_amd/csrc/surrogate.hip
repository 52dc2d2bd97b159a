// placeholder, replaced by the surrogate implementation
#include "qlx_internal.h"
extern "C" {
int qlx_surrogate_train(void*, const qlx_surrogate_cfg*, double*) { return QLX_ERR_STATE; }
int qlx_surrogate_save(void*, const char*) { return QLX_ERR_STATE; }
int qlx_surrogate_load(void*, const char*) { return QLX_ERR_STATE; }
int qlx_surrogate_eval(void*, int, const float*, float*) { return QLX_ERR_STATE; }
int qlx_surrogate_param_count(void*) { return 0; }
int qlx_surrogate_get_params(void*, float*) { return QLX_ERR_STATE; }
int qlx_surrogate_set_params(void*, const float*) { return QLX_ERR_STATE; }
int qlx_surrogate_grad(void*, int, const float*, const float*, float*, float*) { return QLX_ERR_STATE; }
}
