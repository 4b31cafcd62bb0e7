// hp2d_subgrid.hip - the lookup-table subgrid local-inertial model for gfx950:
//   the SFINCS-style tables        Subgrid_Properties_Lookup.m:1-510
//   one routing step               Local_Inertial_Model_D4_Subgrid.m:1-317 (flag_reservoir = 0)
//   a depth change through tables  hp2d_subgrid_apply_volume_change.m (SFINCS-exact branch)
// with the table queries of hp2d_sfincs_{cell_volume_from_zs, zs_from_cell_volume,
// cell_wet_area_from_zs, velocity_state, boundary_velocity_state}.m.
//
// The tables stay resident in HBM beside the handle (one allocation, level
// pages contiguous per field, so a query's level reads of neighbouring cells
// coalesce).  The builder runs one lane per coarse cell or face, walking its
// fine-DEM patch from global memory.  A routing step runs one lane per cell
// for each of: face fluxes (both faces a cell owns, table lookups + the
// semi-implicit friction update), the donor scale of the volume limiter, the
// limited fluxes and candidate volumes (with a device flag for a negative
// volume), and the accepted update (volume -> level -> wet area, outlet
// discharge).  The host reads one flag per internal sub-step to halve or
// accept it, as the reference's adaptive loop does (:52-111).
#include <cfloat>
#include <cmath>
#include <string>
#include <vector>

#include "hp2d.h"
#include "hp2d_handle.hpp"
#include "hp2d_internal.hpp"

namespace hp2d {
namespace {

constexpr int kSgBlock = 256;

struct DevFaces {
    double *zmin, *zmax, *havg, *nrep, *pwet, *navg, *ffit;
};

struct DevTables {
    int ny, nx, nl;
    long long n, nux, nvy;
    double cell_area;
    double *z_zmin, *z_zmax, *z_volmax, *z_level, *z_volume, *z_area;
    DevFaces u, v, side[4];
};

// resident tables of a handle
struct Resident {
    DevTables t{};
    void *mem = nullptr;
    size_t bytes = 0;
};

void resident_free(void *p) {
    Resident *r = (Resident *)p;
    if (r->mem) hipFree(r->mem);
    delete r;
}

long long n_u(int ny, int nx) { return (long long)ny * (nx > 1 ? nx - 1 : 1); }
long long n_v(int ny, int nx) { return (long long)(ny > 1 ? ny - 1 : 1) * nx; }

// the resident tables of a handle, (re)allocated for nl levels
Resident *resident(hp2d_handle *h, int nl, double cell_area, hipError_t &e) {
    int32_t ny = 0, nx = 0, prec = 0;
    hp2d_handle_shape(h, &ny, &nx, &prec);
    HandleSlot *sl = hp2d_handle_slot(h, kSlotSubgrid);
    Resident *r = (Resident *)sl->p;
    const long long n = (long long)ny * nx, nu = n_u(ny, nx), nv = n_v(ny, nx);
    const size_t doubles = (size_t)(3 * n + 3 * n * nl) + (size_t)(nu + nv + 4 * n) * (4 + 3 * nl);
    if (!r) {
        r = new Resident();
        sl->p = r;
        sl->del = resident_free;
    }
    if (r->bytes != doubles * 8) {
        if (r->mem) {
            hipStreamSynchronize(hp2d_handle_stream(h));
            hipFree(r->mem);
            r->mem = nullptr;
            r->bytes = 0;
        }
        e = hipMalloc(&r->mem, doubles * 8);
        if (e != hipSuccess) {
            r->mem = nullptr;
            return nullptr;
        }
        r->bytes = doubles * 8;
    }
    DevTables &t = r->t;
    t.ny = ny;
    t.nx = nx;
    t.nl = nl;
    t.n = n;
    t.nux = nu;
    t.nvy = nv;
    t.cell_area = cell_area;
    double *p = (double *)r->mem;
    auto take = [&](long long k) {
        double *q = p;
        p += k;
        return q;
    };
    t.z_zmin = take(n);
    t.z_zmax = take(n);
    t.z_volmax = take(n);
    t.z_level = take(n * nl);
    t.z_volume = take(n * nl);
    t.z_area = take(n * nl);
    auto faces = [&](DevFaces &f, long long nf) {
        f.zmin = take(nf);
        f.zmax = take(nf);
        f.navg = take(nf);
        f.ffit = take(nf);
        f.havg = take(nf * nl);
        f.nrep = take(nf * nl);
        f.pwet = take(nf * nl);
    };
    faces(t.u, nu);
    faces(t.v, nv);
    for (auto &s : t.side) faces(s, n);
    return r;
}

// ------------------------------------------------------------ table queries --

// the first level interval bracketing x, else the first / last level value
// (hp2d_sfincs_cell_*.m lookup_cell_table_by_z / _by_volume); NaN if none
__device__ double cell_lookup(const double *xl, const double *yl, long long n, long long c, int nl,
                              double x) {
    for (int k = 0; k + 1 < nl; ++k) {
        const double x1 = xl[c + k * n], x2 = xl[c + (k + 1) * n];
        const double y1 = yl[c + k * n], y2 = yl[c + (k + 1) * n];
        if (isfinite(x1) && isfinite(x2) && isfinite(y1) && isfinite(y2) && x2 > x1 && x >= x1 &&
            x <= x2) {
            const double w = (x - x1) / (x2 - x1);
            return y1 + w * (y2 - y1);
        }
    }
    const double xf = xl[c], yf = yl[c];
    const double xe = xl[c + (long long)(nl - 1) * n], ye = yl[c + (long long)(nl - 1) * n];
    if (isfinite(xf) && isfinite(yf) && x <= xf) return yf;
    if (isfinite(xe) && isfinite(ye) && x >= xe) return ye;
    if (isfinite(ye)) return ye;
    return NAN;
}

__device__ double volume_from_zs(const DevTables &T, long long c, double zs) {
    const double zmin = T.z_zmin[c], zmax = T.z_zmax[c];
    double V = 0.0;
    if (isfinite(zs) && isfinite(zmin) && isfinite(zmax)) {
        if (zs <= zmin) V = 0.0;
        else if (zs >= zmax) V = T.z_volmax[c] + (zs - zmax) * T.cell_area;
        else {
            V = cell_lookup(T.z_level, T.z_volume, T.n, c, T.nl, zs);
            if (isnan(V)) V = 0.0;
        }
    }
    if (!isfinite(V)) V = 0.0;
    return fmax(V, 0.0);
}

__device__ double zs_from_volume(const DevTables &T, long long c, double V) {
    const double zmin = T.z_zmin[c], zmax = T.z_zmax[c];
    V = fmax(V, 0.0);
    if (!(isfinite(V) && isfinite(zmin) && isfinite(zmax))) return NAN;
    if (V >= T.z_volmax[c]) return zmax + (V - T.z_volmax[c]) / T.cell_area;
    if (V > 0) return cell_lookup(T.z_volume, T.z_level, T.n, c, T.nl, V);
    return zmin;
}

__device__ double wet_area(const DevTables &T, long long c, double zs) {
    const double zmin = T.z_zmin[c], zmax = T.z_zmax[c];
    double A = 0.0;
    if (isfinite(zs) && isfinite(zmin) && isfinite(zmax)) {
        if (zs >= zmax) A = T.cell_area;
        else if (zs > zmin) {
            A = cell_lookup(T.z_level, T.z_area, T.n, c, T.nl, zs);
            if (isnan(A)) A = 0.0;
        }
    }
    if (!isfinite(A)) A = 0.0;
    return fmax(fmin(A, T.cell_area), 0.0);
}

// lookup_levels: linear interpolation on the uniform level axis zmin..zmax
__device__ double level_lookup(const double *tab, long long nf, long long f, int nl, double zmin,
                               double zmax, double zu) {
    if (zmax <= zmin) return tab[f + (long long)(nl - 1) * nf];
    double pos = (zu - zmin) / (zmax - zmin);
    pos = fmax(fmin(pos, 1.0), 0.0);
    const double fidx = pos * (nl - 1) + 1.0;
    double i1 = floor(fidx);
    i1 = fmax(fmin(i1, (double)(nl - 1)), 1.0);
    const double w = fidx - i1;
    const long long k1 = (long long)i1 - 1;
    const double v1 = tab[f + k1 * nf], v2 = tab[f + (k1 + 1) * nf];
    return v1 + w * (v2 - v1);
}

// query_table_set: face depth HG, representative n, wet fraction phi
__device__ void face_state(const DevFaces &F, long long nf, long long f, int nl, double zu,
                           double &HG, double &nr, double &phi) {
    const double zmin = F.zmin[f], zmax = F.zmax[f];
    double havg = 0.0, nrep = INFINITY, pwet = 0.0;
    const bool valid = isfinite(zu) && isfinite(zmin) && isfinite(zmax);
    if (valid && zu >= zmin && zu < zmax) {
        havg = level_lookup(F.havg, nf, f, nl, zmin, zmax, zu);
        nrep = level_lookup(F.nrep, nf, f, nl, zmin, zmax, zu);
        pwet = level_lookup(F.pwet, nf, f, nl, zmin, zmax, zu);
    }
    if (valid && zu >= zmax) {
        const double hM = F.havg[f + (long long)(nl - 1) * nf];
        double nM = F.nrep[f + (long long)(nl - 1) * nf];
        const double navg = F.navg[f], dz = zu - zmax;
        havg = hM + dz;
        if (!isfinite(nM) || nM <= 0) nM = navg;
        const double denom = F.ffit[f] * dz + 1.0;
        double na = navg - (navg - nM) / denom;
        if (!isfinite(na) || na <= 0) na = navg;
        nrep = na;
        pwet = 1.0;
    }
    if (valid && zu < zmin) {
        havg = 0.0;
        pwet = 0.0;
        nrep = INFINITY;
    }
    if (!isfinite(havg)) havg = 0.0;
    if (!isfinite(pwet)) pwet = 0.0;
    if (!isfinite(nrep) || nrep <= 0) nrep = INFINITY;
    HG = fmax(havg, 0.0);
    nr = nrep;
    phi = fmax(fmin(pwet, 1.0), 0.0);
}

// ------------------------------------------------------------------ builder --

struct BuildArgs {
    const double *dem, *rough;
    int fny, fnx, rough_fine;
    double cellsize, coarse_res;
};

// a rectangle of fine cells [r0, r1) x [c0, c1) of coarse cell (ci, cj)
struct Rect {
    int r0, r1, c0, c1, ci, cj;
};

__device__ Rect patch_of(const BuildArgs &b, int i, int j) {
    Rect p;
    p.r0 = (int)floor(i * b.coarse_res / b.cellsize);
    p.r1 = min((int)floor((i + 1) * b.coarse_res / b.cellsize), b.fny);
    p.c0 = (int)floor(j * b.coarse_res / b.cellsize);
    p.c1 = min((int)floor((j + 1) * b.coarse_res / b.cellsize), b.fnx);
    p.ci = i;
    p.cj = j;
    return p;
}

__device__ double fine_n(const BuildArgs &b, int ny, const Rect &p, int r, int c) {
    const double n = b.rough_fine ? b.rough[r + (long long)c * b.fny] : b.rough[p.ci + (long long)p.cj * ny];
    return (!isfinite(n) || n <= 0) ? NAN : n;
}

// a pool of up to two rectangles whose valid entries (finite z, finite n > 0)
// are visited column-major, the first rectangle first
struct Pool {
    Rect r[2];
    int k;
};

template <typename F>
__device__ void pool_for(const BuildArgs &b, int ny, const Pool &P, F &&f) {
    for (int q = 0; q < P.k; ++q) {
        const Rect &R = P.r[q];
        for (int c = R.c0; c < R.c1; ++c)
            for (int r = R.r0; r < R.r1; ++r) {
                const double z = b.dem[r + (long long)c * b.fny], n = fine_n(b, ny, R, r, c);
                if (isfinite(z) && isfinite(n) && n > 0) f(z, n);
            }
    }
}

__device__ double lin(double d1, double d2, int nl, int k) {
    const int n1 = nl - 1;
    if (d1 == d2 || k == 0) return d1;
    if (k == n1) return d2;
    return d1 + ((double)k * (d2 - d1)) / n1;
}

// sfincs_velocity_table_from_vectors over a pool, written to face f of F
__device__ void face_table(const BuildArgs &b, int ny, const Pool &P, double zmin, double zmax,
                           int nl, const DevFaces &F, long long nf, long long f) {
    double m = 0.0, sn = 0.0;
    pool_for(b, ny, P, [&](double, double n) {
        m += 1.0;
        sn += n;
    });
    if (m == 0.0 || !isfinite(zmin) || !isfinite(zmax)) return;  // stays NaN
    double hM = 0.0, nM = INFINITY;
    for (int k = 0; k < nl; ++k) {
        const double zu = zmax > zmin ? lin(zmin, zmax, nl, k) : zmin;
        double sh = 0.0, sw = 0.0, sd = 0.0;
        pool_for(b, ny, P, [&](double z, double n) {
            const double h = fmax(zu - z, 0.0);
            sw += (double)(h > 0);
            sh += h;
            sd += pow(fmax(zu - fmax(z, zmin), 0.0), 5.0 / 3.0) / n;
        });
        const double havg = sh / m, denom = sd / m;
        const double nrep = (havg > 0 && denom > 0 && isfinite(denom)) ? pow(havg, 5.0 / 3.0) / denom
                                                                      : INFINITY;
        F.havg[f + k * nf] = havg;
        F.nrep[f + k * nf] = nrep;
        F.pwet[f + k * nf] = sw / m;
        hM = havg;
        nM = nrep;
    }
    const double navg = sn / m;
    const double dzfit = fmax(zmax - zmin, DBL_EPSILON), zfit = zmax + dzfit;
    double sf = 0.0;
    pool_for(b, ny, P, [&](double z, double n) {
        sf += pow(fmax(zfit - fmax(z, zmin), 0.0), 5.0 / 3.0) / n;
    });
    const double denom_fit = sf / m;
    const double nfit = (denom_fit > 0 && isfinite(denom_fit))
                             ? pow(hM + zfit - zmax, 5.0 / 3.0) / denom_fit : nM;
    double beta = 0.0;
    if (isfinite(navg) && isfinite(nM) && isfinite(nfit) && fabs(navg - nfit) > DBL_EPSILON)
        beta = ((navg - nM) / (navg - nfit) - 1.0) / dzfit;
    if (!isfinite(beta)) beta = 0.0;
    F.zmin[f] = zmin;
    F.zmax[f] = zmax;
    F.navg[f] = navg;
    F.ffit[f] = beta;
}

__device__ void pool_minmax(const BuildArgs &b, int ny, const Pool &P, double &lo, double &hi,
                            double &cnt) {
    lo = INFINITY;
    hi = -INFINITY;
    cnt = 0.0;
    pool_for(b, ny, P, [&](double z, double) {
        lo = fmin(lo, z);
        hi = fmax(hi, z);
        cnt += 1.0;
    });
}

__device__ int half_len(int len) {
    const int h = (int)round(len / 2.0);
    return h < 1 ? 1 : h;
}

// fill every table value with NaN (the reference's initial values)
__global__ __launch_bounds__(kSgBlock) void sg_nan_kernel(double *p, long long count) {
    for (long long k = blockIdx.x * (long long)kSgBlock + threadIdx.x; k < count;
         k += (long long)gridDim.x * kSgBlock)
        p[k] = NAN;
}

// cell tables (sfincs_cell_table, :186-207) and the four boundary tables
__global__ __launch_bounds__(kSgBlock) void sg_cell_build_kernel(BuildArgs b, DevTables T) {
    const long long c = blockIdx.x * (long long)kSgBlock + threadIdx.x;
    if (c >= T.n) return;
    const int ny = T.ny, i = (int)(c % ny), j = (int)(c / ny), nl = T.nl;
    const Rect p = patch_of(b, i, j);
    if (p.r1 <= p.r0 || p.c1 <= p.c0) return;
    const double fa = b.cellsize * b.cellsize;
    double zmin = INFINITY, zmax = -INFINITY, m = 0.0;
    for (int cc = p.c0; cc < p.c1; ++cc)
        for (int r = p.r0; r < p.r1; ++r) {
            const double z = b.dem[r + (long long)cc * b.fny];
            if (isfinite(z)) {
                zmin = fmin(zmin, z);
                zmax = fmax(zmax, z);
                m += 1.0;
            }
        }
    if (m == 0.0) return;
    auto vol = [&](double zs) {
        double s = 0.0;
        for (int cc = p.c0; cc < p.c1; ++cc)
            for (int r = p.r0; r < p.r1; ++r) {
                const double z = b.dem[r + (long long)cc * b.fny];
                if (isfinite(z)) s += fmax(zs - z, 0.0);
            }
        return s * fa;
    };
    const double volmax = vol(zmax);
    T.z_zmin[c] = zmin;
    T.z_zmax[c] = zmax;
    T.z_volmax[c] = volmax;
    for (int k = 0; k < nl; ++k) {
        const double Vk = lin(0.0, volmax, nl, k);
        double zl;  // invert_volume_exact (:208-231)
        if (!isfinite(Vk) || Vk <= 0) zl = zmin;
        else if (volmax <= DBL_EPSILON || zmax <= zmin) zl = zmax + Vk / (m * fa);
        else if (Vk >= volmax) zl = zmax + (Vk - volmax) / (m * fa);
        else {
            double lo = zmin, hi = zmax;
            for (int it = 0; it < 60; ++it) {
                const double mid = 0.5 * (lo + hi);
                if (vol(mid) < Vk) lo = mid;
                else hi = mid;
            }
            zl = 0.5 * (lo + hi);
        }
        double wet = 0.0;
        for (int cc = p.c0; cc < p.c1; ++cc)
            for (int r = p.r0; r < p.r1; ++r) {
                const double z = b.dem[r + (long long)cc * b.fny];
                if (isfinite(z)) wet += (double)(zl > z);
            }
        T.z_level[c + k * T.n] = zl;
        T.z_volume[c + k * T.n] = Vk;
        T.z_area[c + k * T.n] = wet * fa;
    }
    // boundary tables (boundary_velocity_support + sfincs_velocity_table_single)
    const int hr = half_len(p.r1 - p.r0), hc = half_len(p.c1 - p.c0);
    for (int s = 0; s < 4; ++s) {
        Pool P;
        P.k = 1;
        P.r[0] = p;
        if (s == 0) P.r[0].r1 = p.r0 + hr;        // north
        else if (s == 1) P.r[0].r0 = p.r1 - hr;   // south
        else if (s == 2) P.r[0].c1 = p.c0 + hc;   // west
        else P.r[0].c0 = p.c1 - hc;               // east
        double lo, hi, cnt;
        pool_minmax(b, ny, P, lo, hi, cnt);
        if (cnt == 0.0) continue;
        face_table(b, ny, P, lo, hi, nl, T.side[s], T.n, c);
    }
}

// velocity-point tables of the x faces (u, dir = 0) or y faces (v, dir = 1)
__global__ __launch_bounds__(kSgBlock) void sg_face_build_kernel(BuildArgs b, DevTables T, int dir) {
    const long long f = blockIdx.x * (long long)kSgBlock + threadIdx.x;
    const int ny = T.ny, nx = T.nx;
    int i, j;
    if (dir == 0) {
        if (nx < 2 || f >= (long long)ny * (nx - 1)) return;
        i = (int)(f % ny);
        j = (int)(f / ny);
    } else {
        if (ny < 2 || f >= (long long)(ny - 1) * nx) return;
        i = (int)(f % (ny - 1));
        j = (int)(f / (ny - 1));
    }
    const Rect A = patch_of(b, i, j), B = dir == 0 ? patch_of(b, i, j + 1) : patch_of(b, i + 1, j);
    if (A.c1 <= A.c0 || B.c1 <= B.c0 || A.r1 <= A.r0 || B.r1 <= B.r0) return;
    Pool PA, PB, P;
    PA.k = PB.k = 1;
    PA.r[0] = A;
    PB.r[0] = B;
    if (dir == 0) {  // east half of A, west half of B
        PA.r[0].c0 = A.c1 - half_len(A.c1 - A.c0);
        PB.r[0].c1 = B.c0 + half_len(B.c1 - B.c0);
    } else {         // south half of A, north half of B
        PA.r[0].r0 = A.r1 - half_len(A.r1 - A.r0);
        PB.r[0].r1 = B.r0 + half_len(B.r1 - B.r0);
    }
    double loA, hiA, cA, loB, hiB, cB;
    pool_minmax(b, ny, PA, loA, hiA, cA);
    pool_minmax(b, ny, PB, loB, hiB, cB);
    if (cA == 0.0 || cB == 0.0) return;
    P.k = 2;
    P.r[0] = PA.r[0];
    P.r[1] = PB.r[0];
    face_table(b, ny, P, fmax(loA, loB), fmax(hiA, hiB), T.nl, dir == 0 ? T.u : T.v,
               dir == 0 ? T.nux : T.nvy, f);
}

// -------------------------------------------------------------------- step --

struct StepArgs {
    DevTables T;
    double R, dt, g;
    const uint8_t *inactive;  // idx_nan or NULL
    const int8_t *side;       // outlet side per cell (-1: not an outlet)
    int outlet_type;
    double slope_outlet;
    double *eta, *V, *Vc, *Ca;
    double *qx, *qy, *qxc, *qyc, *qxl, *qyl, *Hxc, *Hyc, *scale;
    double *outvol, *Hf;
    int *negative;
};

__device__ double lie(double qold, double HG, double nr, double slope, double dt, double g,
                      bool active) {
    if (!active) return 0.0;
    const double den = 1.0 + g * dt * (nr * nr) * fabs(qold) / fmax(pow(HG, 7.0 / 3.0), DBL_EPSILON);
    const double rhs = qold - g * dt * HG * slope;
    const double q = rhs / den;
    return isfinite(q) ? q : 0.0;
}

// sfincs_flux_update (:168-208): the east and south faces a cell owns
__global__ __launch_bounds__(kSgBlock) void sg_flux_kernel(StepArgs a) {
    const long long c = blockIdx.x * (long long)kSgBlock + threadIdx.x;
    const DevTables &T = a.T;
    if (c >= T.n) return;
    const int ny = T.ny, i = (int)(c % ny), j = (int)(c / ny);
    double qx = 0.0, hx = 0.0, qy = 0.0, hy = 0.0;
    if (j + 1 < T.nx) {
        const long long e = c + ny;
        const double zu = fmax(a.eta[c], a.eta[e]);
        double HG, nr, phi;
        face_state(T.u, T.nux, c, T.nl, zu, HG, nr, phi);
        const double slope = (a.eta[e] - a.eta[c]) / a.R;
        const bool inact = a.inactive && (a.inactive[c] || a.inactive[e]);
        const bool active = !inact && HG > 0 && phi > 0 && isfinite(nr) && nr > 0 && isfinite(slope);
        qx = lie(a.qx[c], HG, nr, slope, a.dt, a.g, active);
        hx = HG;
    }
    if (i + 1 < ny) {
        const long long s = c + 1, f = i + (long long)j * (ny - 1);
        const double zu = fmax(a.eta[c], a.eta[s]);
        double HG, nr, phi;
        face_state(T.v, T.nvy, f, T.nl, zu, HG, nr, phi);
        const double slope = (a.eta[s] - a.eta[c]) / a.R;
        const bool inact = a.inactive && (a.inactive[c] || a.inactive[s]);
        const bool active = !inact && HG > 0 && phi > 0 && isfinite(nr) && nr > 0 && isfinite(slope);
        qy = lie(a.qy[c], HG, nr, slope, a.dt, a.g, active);
        hy = HG;
    }
    a.qxc[c] = qx;
    a.qyc[c] = qy;
    a.Hxc[c] = hx;
    a.Hyc[c] = hy;
}

// limit_face_discharge_by_available_volume (:132-166): the donor scale
__global__ __launch_bounds__(kSgBlock) void sg_scale_kernel(StepArgs a) {
    const long long c = blockIdx.x * (long long)kSgBlock + threadIdx.x;
    const DevTables &T = a.T;
    if (c >= T.n) return;
    const int ny = T.ny, i = (int)(c % ny), j = (int)(c / ny);
    const double Qw = j > 0 ? a.qxc[c - ny] * a.R : 0.0, Qe = a.qxc[c] * a.R;
    const double Qn = i > 0 ? a.qyc[c - 1] * a.R : 0.0, Qs = a.qyc[c] * a.R;
    const double out = fmax(Qe, 0.0) + fmax(-Qw, 0.0) + fmax(Qs, 0.0) + fmax(-Qn, 0.0);
    double s = 1.0;
    const double V = fmax(a.V[c], 0.0);
    if (out * a.dt > V && out > 0) s = V / (out * a.dt);
    a.scale[c] = fmax(fmin(s, 1.0), 0.0);
}

// the limited discharge of the x face east of cell c (y face south of c)
__device__ double limited_x(const StepArgs &a, long long c, int j) {
    const double Q = a.qxc[c] * a.R;
    if (j + 1 >= a.T.nx) return Q;
    return Q * (Q >= 0 ? a.scale[c] : a.scale[c + a.T.ny]);
}
__device__ double limited_y(const StepArgs &a, long long c, int i) {
    const double Q = a.qyc[c] * a.R;
    if (i + 1 >= a.T.ny) return Q;
    return Q * (Q >= 0 ? a.scale[c] : a.scale[c + 1]);
}

// limited fluxes and the candidate volume (:62-82)
__global__ __launch_bounds__(kSgBlock) void sg_volume_kernel(StepArgs a) {
    const long long c = blockIdx.x * (long long)kSgBlock + threadIdx.x;
    const DevTables &T = a.T;
    if (c >= T.n) return;
    const int ny = T.ny, i = (int)(c % ny), j = (int)(c / ny);
    const double Qe = limited_x(a, c, j), Qs = limited_y(a, c, i);
    const double Qw = j > 0 ? limited_x(a, c - ny, j - 1) : 0.0;
    const double Qn = i > 0 ? limited_y(a, c - 1, i - 1) : 0.0;
    a.qxl[c] = Qe / a.R;
    a.qyl[c] = Qs / a.R;
    const double Vc = a.V[c] + a.dt * (Qw - Qe + Qn - Qs);
    a.Vc[c] = Vc;
    if (Vc < -1e-9) atomicOr(a.negative, 1);
}

// the accepted sub-step (:84-107, apply_sfincs_outlet :242-284)
__global__ __launch_bounds__(kSgBlock) void sg_accept_kernel(StepArgs a) {
    const long long c = blockIdx.x * (long long)kSgBlock + threadIdx.x;
    const DevTables &T = a.T;
    if (c >= T.n) return;
    double V = fmax(a.Vc[c], 0.0);
    double eta = zs_from_volume(T, c, V);
    double Ca = wet_area(T, c, eta);
    double oflow = 0.0, Hout = 0.0;
    const int s = a.side ? a.side[c] : -1;
    if (s >= 0) {
        double HG, nr, phi, Qo = 0.0;
        face_state(T.side[s], T.n, c, T.nl, eta, HG, nr, phi);
        if (HG > 0 && isfinite(nr) && nr > 0) {
            const double qg = a.outlet_type == 1
                                  ? (1.0 / nr) * pow(HG, 5.0 / 3.0) * sqrt(fmax(fabs(a.slope_outlet), 0.0))
                                  : HG * sqrt(a.g * fmax(HG, 0.0));
            Qo = qg * a.R;
            Hout = HG;
        }
        Qo = fmin(fmax(Qo, 0.0), fmax(V, 0.0) / a.dt);
        oflow = Qo / (a.R * a.R) * 1000.0 * 3600.0;
        V = fmax(V - Qo * a.dt, 0.0);
        eta = zs_from_volume(T, c, V);
        Ca = wet_area(T, c, eta);
    }
    a.V[c] = V;
    a.eta[c] = eta;
    a.Ca[c] = Ca;
    a.outvol[c] = a.outvol[c] + oflow / 1000.0 / 3600.0 * (a.R * a.R) * a.dt;
    a.Hf[c] = a.Hxc[c];
    a.Hf[T.n + c] = a.Hyc[c];
    a.Hf[2 * T.n + c] = Hout;
    a.qx[c] = a.qxl[c];
    a.qy[c] = a.qyl[c];
}

// initial state (:28-43)
__global__ __launch_bounds__(kSgBlock) void sg_init_kernel(StepArgs a, const double *d_tot,
                                                           const double *prev, int n_prev) {
    const long long c = blockIdx.x * (long long)kSgBlock + threadIdx.x;
    const DevTables &T = a.T;
    if (c >= T.n) return;
    const double eta = T.z_zmin[c] + fmax(d_tot[c] / 1000.0, 0.0);
    a.eta[c] = eta;
    a.V[c] = volume_from_zs(T, c, eta);
    a.Ca[c] = wet_area(T, c, eta);
    a.qx[c] = n_prev >= 1 ? prev[c] / 1000.0 / 3600.0 * a.R : 0.0;
    a.qy[c] = n_prev >= 2 ? prev[T.n + c] / 1000.0 / 3600.0 * a.R : 0.0;
    a.outvol[c] = 0.0;
    a.Hf[c] = a.Hf[T.n + c] = a.Hf[2 * T.n + c] = 0.0;
}

// the outputs (:113-124): outflow pages, d_t, I_tot_end_cell
__global__ __launch_bounds__(kSgBlock) void sg_final_kernel(StepArgs a, double dt_total,
                                                            double *outflow, double *d_t,
                                                            double *Itot) {
    const long long c = blockIdx.x * (long long)kSgBlock + threadIdx.x;
    const DevTables &T = a.T;
    if (c >= T.n) return;
    const double R = a.R;
    const double o1 = a.qx[c] / R * 1000.0 * 3600.0, o2 = a.qy[c] / R * 1000.0 * 3600.0;
    const double o3 = a.outvol[c] / (R * R) * 1000.0 * 3600.0 / dt_total;
    outflow[c] = o1;
    outflow[T.n + c] = o2;
    outflow[2 * T.n + c] = o3;
    d_t[c] = fmax(a.eta[c] - T.z_zmin[c], 0.0) * 1000.0;
    Itot[c] = fabs(o1 + o2 + o3) * dt_total / 1000.0 / 3600.0 * (R * R);
}

// hp2d_subgrid_apply_volume_change (SFINCS-exact branch)
__global__ __launch_bounds__(kSgBlock) void sg_volume_change_kernel(DevTables T, const double *d0,
                                                                    const double *delta,
                                                                    double area, double *d1,
                                                                    double *Vo, double *Vn,
                                                                    double *A, double *res) {
    const long long c = blockIdx.x * (long long)kSgBlock + threadIdx.x;
    if (c >= T.n) return;
    const double dv = delta[c] / 1000.0 * area;
    const double vo = volume_from_zs(T, c, T.z_zmin[c] + fmax(d0[c], 0.0));
    const double vn = fmax(vo + dv, 0.0);
    const double zs = zs_from_volume(T, c, vn);
    d1[c] = fmax(zs - T.z_zmin[c], 0.0);
    Vo[c] = vo;
    Vn[c] = vn;
    A[c] = wet_area(T, c, zs);
    res[c] = (vo + dv < 0) ? vn : vn - vo - dv;
}

int grid_of(long long n) { return (int)((n + kSgBlock - 1) / kSgBlock); }

// host <-> device copies of every table array
template <typename Fn>
void each_array(const hp2d_subgrid_tables *h, DevTables &d, long long n, long long nu, long long nv,
                int nl, Fn &&fn) {
    fn(h->z_zmin, d.z_zmin, n);
    fn(h->z_zmax, d.z_zmax, n);
    fn(h->z_volmax, d.z_volmax, n);
    fn(h->z_level, d.z_level, n * nl);
    fn(h->z_volume, d.z_volume, n * nl);
    fn(h->z_area, d.z_area, n * nl);
    auto faces = [&](const hp2d_subgrid_faces &hf, DevFaces &df, long long nf) {
        fn(hf.zmin, df.zmin, nf);
        fn(hf.zmax, df.zmax, nf);
        fn(hf.navg, df.navg, nf);
        fn(hf.ffit, df.ffit, nf);
        fn(hf.havg, df.havg, nf * nl);
        fn(hf.nrep, df.nrep, nf * nl);
        fn(hf.pwet, df.pwet, nf * nl);
    };
    faces(h->u, d.u, nu);
    faces(h->v, d.v, nv);
    for (int s = 0; s < 4; ++s) faces(h->side[s], d.side[s], n);
}

bool tables_complete(const hp2d_subgrid_tables *t) {
    if (!t->z_zmin || !t->z_zmax || !t->z_volmax || !t->z_level || !t->z_volume || !t->z_area)
        return false;
    auto ok = [](const hp2d_subgrid_faces &f) {
        return f.zmin && f.zmax && f.havg && f.nrep && f.pwet && f.navg && f.ffit;
    };
    return ok(t->u) && ok(t->v) && ok(t->side[0]) && ok(t->side[1]) && ok(t->side[2]) &&
           ok(t->side[3]);
}

Resident *resident_of(hp2d_handle *h) {
    HandleSlot *sl = hp2d_handle_slot(h, kSlotSubgrid);
    Resident *r = (Resident *)sl->p;
    return (r && r->mem) ? r : nullptr;
}

}  // namespace
}  // namespace hp2d

using namespace hp2d;

extern "C" hp2d_status hp2d_subgrid_tables_build(hp2d_handle *h, const double *dem, int32_t fny,
                                                 int32_t fnx, double cellsize,
                                                 const double *roughness, int32_t rough_is_fine,
                                                 double coarse_res, int32_t nr_levels,
                                                 hp2d_subgrid_tables *out) {
    if (!h || !dem || !roughness || fny < 1 || fnx < 1 || !(cellsize > 0) || !(coarse_res > 0) ||
        nr_levels < 2 || nr_levels > 256)
        return HP2D_E_ARG;
    const double r = coarse_res / cellsize;
    if (std::lround(r) % 2 != 0)
        return hp2d_handle_fail(h, "hp2d_subgrid_tables_build: the coarse-to-fine ratio must round "
                                   "to an even integer (velocity-point tables take half cells)");
    if (out && !tables_complete(out)) return HP2D_E_ARG;
    int32_t ny = 0, nx = 0, prec = 0;
    hp2d_handle_shape(h, &ny, &nx, &prec);
    const size_t nf = (size_t)fny * fnx, nr = rough_is_fine ? nf : (size_t)ny * nx;
    hipStream_t s = hp2d_handle_stream(h);
    hipError_t e = hipSuccess;
    Resident *res = resident(h, nr_levels, coarse_res * coarse_res, e);
    if (!res) return hp2d_handle_fail(h, std::string("hp2d_subgrid_tables_build: ") + hipGetErrorString(e));
    double *fine = (double *)hp2d_handle_workspace(h, (nf + nr) * 8 + 256);
    if (!fine) return HP2D_E_NOMEM;
    e = hipMemcpyAsync(fine, dem, nf * 8, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(fine + nf, roughness, nr * 8, hipMemcpyHostToDevice, s);
    DevTables T = res->t;
    const long long total = (long long)(res->bytes / 8);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(sg_nan_kernel, dim3(std::min<long long>(grid_of(total), 65536)),
                           dim3(kSgBlock), 0, s, (double *)res->mem, total);
        e = hipGetLastError();
    }
    BuildArgs b{fine, fine + nf, fny, fnx, rough_is_fine ? 1 : 0, cellsize, coarse_res};
    if (e == hipSuccess) {
        hipLaunchKernelGGL(sg_cell_build_kernel, dim3(grid_of(T.n)), dim3(kSgBlock), 0, s, b, T);
        e = hipGetLastError();
    }
    if (e == hipSuccess && nx > 1) {
        hipLaunchKernelGGL(sg_face_build_kernel, dim3(grid_of((long long)ny * (nx - 1))),
                           dim3(kSgBlock), 0, s, b, T, 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess && ny > 1) {
        hipLaunchKernelGGL(sg_face_build_kernel, dim3(grid_of((long long)(ny - 1) * nx)),
                           dim3(kSgBlock), 0, s, b, T, 1);
        e = hipGetLastError();
    }
    if (out && e == hipSuccess) {
        out->nr_levels = nr_levels;
        out->cell_area = coarse_res * coarse_res;
        each_array(out, T, T.n, T.nux, T.nvy, nr_levels, [&](double *hp, double *dp, long long cnt) {
            if (e == hipSuccess) e = hipMemcpyAsync(hp, dp, cnt * 8, hipMemcpyDeviceToHost, s);
        });
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return hp2d_handle_fail(h, std::string("hp2d_subgrid_tables_build: ") + hipGetErrorString(e));
    return HP2D_OK;
}

extern "C" hp2d_status hp2d_subgrid_set_tables(hp2d_handle *h, const hp2d_subgrid_tables *t) {
    if (!h || !t || t->nr_levels < 2 || t->nr_levels > 256 || !(t->cell_area > 0) ||
        !tables_complete(t))
        return HP2D_E_ARG;
    hipError_t e = hipSuccess;
    Resident *res = resident(h, t->nr_levels, t->cell_area, e);
    if (!res) return hp2d_handle_fail(h, std::string("hp2d_subgrid_set_tables: ") + hipGetErrorString(e));
    hipStream_t s = hp2d_handle_stream(h);
    DevTables &T = res->t;
    each_array(t, T, T.n, T.nux, T.nvy, T.nl, [&](double *hp, double *dp, long long cnt) {
        if (e == hipSuccess) e = hipMemcpyAsync(dp, hp, cnt * 8, hipMemcpyHostToDevice, s);
    });
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return hp2d_handle_fail(h, std::string("hp2d_subgrid_set_tables: ") + hipGetErrorString(e));
    return HP2D_OK;
}

extern "C" hp2d_status hp2d_local_inertial_d4_subgrid(
    hp2d_handle *h, const double *d_tot, double time_step_min, double resolution,
    int32_t outlet_type, double slope_outlet, const int32_t *row_outlet, const int32_t *col_outlet,
    int32_t n_outlet, const double *outflow_prev, int32_t n_prev_pages, const uint8_t *idx_nan,
    double *outflow, double *d_t, double *I_tot_end_cell, double *Hf, double *C_a, double *eta_t,
    double *V_t, int32_t *n_substeps) {
    if (!h || !d_tot || n_outlet < 0 || (n_outlet > 0 && (!row_outlet || !col_outlet)) ||
        n_prev_pages < 0 || (n_prev_pages > 0 && !outflow_prev) || !(resolution > 0) ||
        !(time_step_min >= 0))
        return HP2D_E_ARG;
    Resident *res = resident_of(h);
    if (!res)
        return hp2d_handle_fail(h, "hp2d_local_inertial_d4_subgrid: no subgrid tables on this "
                                   "handle (hp2d_subgrid_tables_build or hp2d_subgrid_set_tables)");
    const DevTables &T = res->t;
    const int ny = T.ny, nx = T.nx;
    const long long n = T.n;
    // outlet sides (outlet_boundary_side, :286-308)
    std::vector<int8_t> side((size_t)n, -1);
    for (int k = 0; k < n_outlet; ++k) {
        const int r = row_outlet[k], c = col_outlet[k];
        if (r < 1 || r > ny || c < 1 || c > nx)
            return hp2d_handle_fail(h, "hp2d_local_inertial_d4_subgrid: outlet row/col off the grid");
        int sd;
        if (r <= 1) sd = 0;
        else if (r >= ny) sd = 1;
        else if (c <= 1) sd = 2;
        else if (c >= nx) sd = 3;
        else {
            const int dist[4] = {r - 1, ny - r, c - 1, nx - c};
            sd = 0;
            for (int q = 1; q < 4; ++q)
                if (dist[q] < dist[sd]) sd = q;
        }
        side[(size_t)(r - 1) + (size_t)(c - 1) * ny] = (int8_t)sd;
    }
    const int n_prev = n_prev_pages < 2 ? n_prev_pages : 2;
    // workspace: 18 arrays + 3 Hf pages + outflow 3 pages | d_tot, prev pages | side, mask | flag
    const size_t nb = (size_t)n * 8;
    const size_t n_dbl = 24 + 1 + (size_t)n_prev;
    char *ws = (char *)hp2d_handle_workspace(h, n_dbl * nb + 2 * (size_t)n + 512);
    if (!ws) return HP2D_E_NOMEM;
    double *w = (double *)ws;
    auto take = [&](size_t k) {
        double *q = w;
        w += k * n;
        return q;
    };
    StepArgs a;
    a.T = T;
    a.R = resolution;
    a.g = 9.81;
    a.outlet_type = outlet_type;
    a.slope_outlet = slope_outlet;
    a.eta = take(1);
    a.V = take(1);
    a.Vc = take(1);
    a.Ca = take(1);
    a.qx = take(1);
    a.qy = take(1);
    a.qxc = take(1);
    a.qyc = take(1);
    a.qxl = take(1);
    a.qyl = take(1);
    a.Hxc = take(1);
    a.Hyc = take(1);
    a.scale = take(1);
    a.outvol = take(1);
    a.Hf = take(3);
    double *dout = take(3), *ddt = take(1), *dI = take(1);
    double *dd = take(1), *dprev = take(n_prev);
    take(2);  // spare
    int8_t *dside = (int8_t *)w;
    uint8_t *dmask = (uint8_t *)(dside + n);
    int *dflag = (int *)(((uintptr_t)(dmask + n) + 255) & ~(uintptr_t)255);
    a.side = n_outlet > 0 ? dside : nullptr;
    a.inactive = idx_nan ? dmask : nullptr;
    a.negative = dflag;
    hipStream_t s = hp2d_handle_stream(h);
    hipError_t e = hipMemcpyAsync(dd, d_tot, nb, hipMemcpyHostToDevice, s);
    for (int p = 0; p < n_prev && e == hipSuccess; ++p)
        e = hipMemcpyAsync(dprev + p * n, outflow_prev + p * n, nb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && n_outlet > 0)
        e = hipMemcpyAsync(dside, side.data(), (size_t)n, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && idx_nan) e = hipMemcpyAsync(dmask, idx_nan, (size_t)n, hipMemcpyHostToDevice, s);
    const int G = grid_of(n);
    a.dt = 0.0;
    if (e == hipSuccess) {
        hipLaunchKernelGGL(sg_init_kernel, dim3(G), dim3(kSgBlock), 0, s, a, dd, dprev, n_prev);
        e = hipGetLastError();
    }
    // the adaptive internal sub-steps (:52-111)
    const double dt_total = time_step_min * 60.0;
    double dt_remaining = dt_total, dt_try = dt_total;
    const double min_sub_dt = std::fmax(dt_total / 64.0, 1e-4);
    int subs = 0;
    long long guard = 0;
    int neg = 0;
    if (e == hipSuccess) e = hp2d_handle_mark(h, 0);
    while (e == hipSuccess && dt_remaining > 1e-12) {
        if (++guard > 1000000) {
            return hp2d_handle_fail(h, "hp2d_local_inertial_d4_subgrid: the sub-step loop does not end");
        }
        const double dt_sub = std::fmin(dt_try, dt_remaining);
        a.dt = dt_sub;
        e = hipMemsetAsync(dflag, 0, sizeof(int), s);
        if (e != hipSuccess) break;
        hipLaunchKernelGGL(sg_flux_kernel, dim3(G), dim3(kSgBlock), 0, s, a);
        hipLaunchKernelGGL(sg_scale_kernel, dim3(G), dim3(kSgBlock), 0, s, a);
        hipLaunchKernelGGL(sg_volume_kernel, dim3(G), dim3(kSgBlock), 0, s, a);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(&neg, dflag, sizeof(int), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) break;
        if (neg && dt_sub > min_sub_dt) {
            dt_try = std::fmax(0.5 * dt_sub, min_sub_dt);
            continue;
        }
        hipLaunchKernelGGL(sg_accept_kernel, dim3(G), dim3(kSgBlock), 0, s, a);
        e = hipGetLastError();
        ++subs;
        dt_remaining = dt_remaining - dt_sub;
        dt_try = std::fmin(dt_try * 1.25, std::fmax(dt_remaining, min_sub_dt));
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(sg_final_kernel, dim3(G), dim3(kSgBlock), 0, s, a, dt_total, dout, ddt, dI);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hp2d_handle_mark(h, 1);
    auto D2H = [&](double *dst, const double *src, size_t bytes) {
        if (e == hipSuccess && dst) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    };
    D2H(outflow, dout, 3 * nb);
    D2H(d_t, ddt, nb);
    D2H(I_tot_end_cell, dI, nb);
    D2H(Hf, a.Hf, 3 * nb);
    D2H(C_a, a.Ca, nb);
    D2H(eta_t, a.eta, nb);
    D2H(V_t, a.V, nb);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return hp2d_handle_fail(h, std::string("hp2d_local_inertial_d4_subgrid: ") + hipGetErrorString(e));
    if (n_substeps) *n_substeps = subs;
    return HP2D_OK;
}

extern "C" hp2d_status hp2d_subgrid_apply_volume_change(hp2d_handle *h, const double *drep_old,
                                                        const double *delta_depth_mm,
                                                        double cell_area, double *drep_new,
                                                        double *V_old, double *V_new,
                                                        double *A_new, double *residual_m3) {
    if (!h || !drep_old || !delta_depth_mm || !drep_new) return HP2D_E_ARG;
    Resident *res = resident_of(h);
    if (!res)
        return hp2d_handle_fail(h, "hp2d_subgrid_apply_volume_change: no subgrid tables on this handle");
    const DevTables &T = res->t;
    const size_t nb = (size_t)T.n * 8;
    double *w = (double *)hp2d_handle_workspace(h, 7 * nb + 256);
    if (!w) return HP2D_E_NOMEM;
    double *d0 = w, *dl = w + T.n, *o[5];
    for (int k = 0; k < 5; ++k) o[k] = w + (2 + k) * T.n;
    hipStream_t s = hp2d_handle_stream(h);
    hipError_t e = hipMemcpyAsync(d0, drep_old, nb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dl, delta_depth_mm, nb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(sg_volume_change_kernel, dim3(grid_of(T.n)), dim3(kSgBlock), 0, s, T, d0,
                           dl, cell_area, o[0], o[1], o[2], o[3], o[4]);
        e = hipGetLastError();
    }
    double *dst[5] = {drep_new, V_old, V_new, A_new, residual_m3};
    for (int k = 0; k < 5 && e == hipSuccess; ++k)
        if (dst[k]) e = hipMemcpyAsync(dst[k], o[k], nb, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return hp2d_handle_fail(h, std::string("hp2d_subgrid_apply_volume_change: ") + hipGetErrorString(e));
    return HP2D_OK;
}
