// bdd.hip -- branchless distance-driven projector (projector type 5).
#include "geom.hpp"

int bdd_prepare(omg_geom*) { return OMG_OK; }
void bdd_release(omg_geom*) {}

int bdd_forward(const omg_geom*, int64_t, int64_t, const float*, float*, hipStream_t) {
    omg::set_error("projector_type 5 is not built yet");
    return OMG_ERR_UNSUPPORTED;
}
int bdd_backward(const omg_geom*, int64_t, int64_t, const float*, float*, float*, hipStream_t) {
    omg::set_error("projector_type 5 is not built yet");
    return OMG_ERR_UNSUPPORTED;
}
