// hp2d_pending.cpp - entry points declared in include/hp2d.h whose kernels
// are not in this build yet.  They fail loudly (HP2D_E_UNSUPPORTED).
#include <string>

#include "hp2d.h"

extern "C" {

hp2d_status hp2d_boussinesq_2d(hp2d_handle *, const double *, const double *, const double *,
                               const double *, const double *, const double *, const uint8_t *,
                               double, double, double, double, double *, double *, double *,
                               int32_t *) {
    return HP2D_E_UNSUPPORTED;
}

}
