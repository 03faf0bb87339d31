// orth.hip -- orthogonal distance-based projector (projector type 2).
#include "geom.hpp"

int orth_prepare(omg_geom*) { return OMG_OK; }
void orth_release(omg_geom*) {}

int orth_forward(const omg_geom*, int64_t, int64_t, const float*, float*, hipStream_t) {
    omg::set_error("projector_type 2 is not built yet");
    return OMG_ERR_UNSUPPORTED;
}
int orth_backward(const omg_geom*, int64_t, int64_t, const float*, float*, float*, hipStream_t) {
    omg::set_error("projector_type 2 is not built yet");
    return OMG_ERR_UNSUPPORTED;
}
