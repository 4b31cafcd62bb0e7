// hp2d_pending.cpp - entry points declared in include/hp2d.h whose kernels
// are not in this build yet.  They fail loudly (HP2D_E_UNSUPPORTED).
#include <string>

#include "hp2d.h"

extern "C" {

hp2d_status hp2d_ca_routing(hp2d_handle *, const double *, const double *, const double *, double,
                            double, const double *, double, const double *, int32_t, double,
                            const int32_t *, const int32_t *, int32_t, int32_t, double *, double *,
                            double *, double *, double *, double *, double *) {
    return HP2D_E_UNSUPPORTED;
}

hp2d_status hp2d_build_up_wash_off(hp2d_handle *, const double *, const double *, const double *,
                                   const double *, const double *, const double *, const double *,
                                   const double *, double, double, const uint8_t *, int32_t,
                                   const double *, double, double, double, double *, double *,
                                   double *, double *, double *, double *, double *) {
    return HP2D_E_UNSUPPORTED;
}

hp2d_status hp2d_boussinesq_2d(hp2d_handle *, const double *, const double *, const double *,
                               const double *, const double *, const double *, const uint8_t *,
                               double, double, double, double, double *, double *, double *,
                               int32_t *) {
    return HP2D_E_UNSUPPORTED;
}

}
