// Adam, Nesterov and online L-BFGS minimizers (placeholders until their
// device passes land; they fail loudly instead of falling back to the CPU).
#include <hip/hip_runtime.h>
#include "../../include/qmin.h"

struct qm_adam { int dummy; };
struct qm_nesterov { int dummy; };
struct qm_lolbfgs { int dummy; };

extern "C" {
int qm_adam_create(qm_adam** out) { *out = new qm_adam{}; return QM_OK; }
int qm_adam_destroy(qm_adam* m) { delete m; return QM_OK; }
int qm_adam_set_parameters(qm_adam*, double, double, double, double, double) { return QM_ESTATE; }
int qm_adam_set_max_iterations(qm_adam*, int) { return QM_ESTATE; }
int qm_adam_minimize(qm_adam*, qm_ldg*, qm_lattice*) { return QM_ESTATE; }
int qm_adam_get_state(const qm_adam*, int*, double*) { return QM_ESTATE; }
int qm_nesterov_create(qm_nesterov** out) { *out = new qm_nesterov{}; return QM_OK; }
int qm_nesterov_destroy(qm_nesterov* m) { delete m; return QM_OK; }
int qm_nesterov_set_parameters(qm_nesterov*, double, double, double, int) { return QM_ESTATE; }
int qm_nesterov_set_max_iterations(qm_nesterov*, int) { return QM_ESTATE; }
int qm_nesterov_minimize(qm_nesterov*, qm_ldg*, qm_lattice*) { return QM_ESTATE; }
int qm_nesterov_get_state(const qm_nesterov*, int*, double*) { return QM_ESTATE; }
int qm_lolbfgs_create(qm_lolbfgs** out) { *out = new qm_lolbfgs{}; return QM_OK; }
int qm_lolbfgs_destroy(qm_lolbfgs* m) { delete m; return QM_OK; }
int qm_lolbfgs_set_parameters(qm_lolbfgs*, int, double, double, double, double, int) { return QM_ESTATE; }
int qm_lolbfgs_set_max_iterations(qm_lolbfgs*, int) { return QM_ESTATE; }
int qm_lolbfgs_minimize(qm_lolbfgs*, qm_ldg*, qm_lattice*) { return QM_ESTATE; }
int qm_lolbfgs_get_state(const qm_lolbfgs*, int*, double*) { return QM_ESTATE; }
}
