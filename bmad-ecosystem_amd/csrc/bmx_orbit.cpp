// Closed orbits and one-turn matrices of many lattices at once (bmad/code/closed_orbit_calc.f90:192).
//
// The reference finds the fixed point of the one-turn map with a Levenberg-Marquardt search
// (super_mrqmin) whose Jacobian is the one-turn transfer matrix built from every element's
// analytic mat6 (transfer_matrix_calc). Here the one-turn map is the GPU tracking itself: each
// lattice gets 13 particles per iteration, the trial orbit and a +-h displacement in each of the
// six coordinates, all lattices in one batch (one captured launch sequence per iteration). The
// central differences give the one-turn matrix (accurate to about 1e-9 relative: the rounding of
// the tracked coordinates over the step, and the step squared times the map's third
// derivatives), and a Newton step on
// (M - 1) d = -(f(x) - x) moves the orbit; the same convergence test as the reference ends it.
// n_dim = 4 closes x, px, y, py at the given pz with the rf cavities off (their voltage set to
// zero, as closed_orbit_calc turns them off); n_dim = 6 closes all six with the rf on.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "bmx_internal.hpp"

using namespace bmx;

namespace {

constexpr int kProbe = 13;  // trial orbit + 6 x (+h, -h)

// Finite-difference steps: large enough that the rounding of the tracked coordinates (about
// 1e-15 after a turn) divided by the step stays near 1e-9, small enough that the truncation
// error (the step squared times the map's third derivatives) does too.
constexpr double kStep[6] = {1e-6, 1e-6, 1e-6, 1e-6, 1e-5, 1e-6};

// Gaussian elimination with partial pivoting: solves a x = b in place (a is n x n, row-major).
bool solve(int n, double *a, double *b)
{
    for (int c = 0; c < n; c++) {
        int p = c;
        for (int r = c + 1; r < n; r++)
            if (std::fabs(a[r * n + c]) > std::fabs(a[p * n + c])) p = r;
        if (a[p * n + c] == 0) return false;
        if (p != c) {
            for (int k = 0; k < n; k++) std::swap(a[p * n + k], a[c * n + k]);
            std::swap(b[p], b[c]);
        }
        for (int r = c + 1; r < n; r++) {
            const double f = a[r * n + c] / a[c * n + c];
            for (int k = c; k < n; k++) a[r * n + k] -= f * a[c * n + k];
            b[r] -= f * b[c];
        }
    }
    for (int c = n - 1; c >= 0; c--) {
        double s = b[c];
        for (int k = c + 1; k < n; k++) s -= a[c * n + k] * b[k];
        b[c] = s / a[c * n + c];
    }
    return true;
}

// coord_struct%beta as init_coord sets it (init_coord.f90:205-207, convert_pc_to.f90:37-41).
double init_beta(double pz, const Species &sp)
{
    if (pz == 0) return sp.beta0;
    const double pc = sp.p0c * (1 + pz);
    return pc / std::sqrt(pc * pc + sp.mc2_sq);
}

struct Lattices {
    std::vector<bmx_lattice *> owned;  // rf-off copies (n_dim = 4)
    ~Lattices()
    {
        for (bmx_lattice *L : owned) bmx_lattice_destroy(L);
    }
};

}  // namespace

extern "C" {

int32_t bmx_closed_orbit(bmx_ctx *ctx, bmx_lattice *const *lats, int32_t n_lat, int32_t n_dim,
                         double *orbit, double *mat, int32_t max_iter, double rel_tol,
                         double abs_tol, int32_t *status)
{
    if (!ctx || n_lat < 0 || (n_lat > 0 && (!lats || !orbit || !status)) ||
        (n_dim != 4 && n_dim != 6) || max_iter < 1 || !(rel_tol >= 0) || !(abs_tol >= 0))
        return set_error(BMX_ERR_ARG, "bmx_closed_orbit: bad argument");
    if (n_lat == 0) return BMX_OK;
    for (int32_t i = 0; i < n_lat; i++)
        if (!lats[i] || lats[i]->device != ctx->device)
            return set_error(BMX_ERR_ARG, "bmx_closed_orbit: lattice missing or on another device");
    // the tracked lattices: the rf cavities switched off for n_dim = 4
    Lattices held;
    std::vector<bmx_lattice *> use(lats, lats + n_lat);
    if (n_dim == 4) {
        for (int32_t i = 0; i < n_lat; i++) {
            HostTable t = lats[i]->host;
            bool has_rf = false;
            for (const EleHdr &h : t.hdr)
                if (h.kind == K_RF) {
                    const int32_t q = h.off + ((h.flags & F_OFFSET) ? kOffsetLen : 0) +
                                      ((h.flags & (F_APE_ENT | F_APE_EXIT | F_APE_ENT_SAME)) ? 4 : 0);
                    t.prm[q + 1] = 0;  // voltage
                    has_rf = true;
                }
            if (!has_rf) continue;
            bmx_lattice *L = nullptr;
            const int32_t rc = lattice_upload(ctx, std::move(t), &L);
            if (rc != BMX_OK) return rc;
            held.owned.push_back(L);
            use[i] = L;
        }
    }
    const int64_t n = (int64_t)n_lat * kProbe;
    bmx_beam *beam = nullptr;
    int32_t rc = bmx_beam_create(ctx, n, &beam);
    if (rc != BMX_OK) return rc;
    std::vector<bmx_group> groups(n_lat);
    for (int32_t i = 0; i < n_lat; i++) groups[i] = bmx_group{use[i], beam, (int64_t)i * kProbe, kProbe};
    bmx_batch *batch = nullptr;
    rc = bmx_batch_create(ctx, groups.data(), n_lat, &batch);
    if (rc != BMX_OK) {
        bmx_beam_destroy(beam);
        return rc;
    }
    // host SoA of the probes
    std::vector<double> c[7];
    for (auto &v : c) v.assign(n, 0.0);
    std::vector<int32_t> st(n), le(n), lt(n);
    bmx_beam_view view{c[0].data(), c[1].data(), c[2].data(), c[3].data(), c[4].data(), c[5].data(),
                       c[6].data(), st.data(), le.data(), lt.data()};
    std::vector<double> x(orbit, orbit + 6 * (size_t)n_lat);
    std::vector<double> jac(36 * (size_t)n_lat, 0.0);
    std::vector<char> done(n_lat, 0);
    for (int32_t i = 0; i < n_lat; i++) status[i] = 1;  // not converged (yet)
    int32_t active = n_lat;
    for (int32_t it = 0; it < max_iter && active > 0 && rc == BMX_OK; it++) {
        for (int32_t i = 0; i < n_lat; i++) {
            for (int k = 0; k < kProbe; k++) {
                const int64_t j = (int64_t)i * kProbe + k;
                double v[6];
                for (int d = 0; d < 6; d++) v[d] = x[6 * i + d];
                if (k > 0) v[(k - 1) / 2] += (k % 2 ? 1 : -1) * kStep[(k - 1) / 2];
                for (int d = 0; d < 6; d++) c[d][j] = v[d];
                c[6][j] = init_beta(v[5], use[i]->host.sp);
                st[j] = BMX_STATE_ALIVE;
                le[j] = 0;
                lt[j] = 0;
            }
        }
        rc = bmx_beam_upload(beam, &view, n);
        if (rc == BMX_OK) rc = bmx_batch_track(batch, 1, 0);
        if (rc == BMX_OK) rc = bmx_beam_download(beam, &view, n);
        if (rc != BMX_OK) break;
        for (int32_t i = 0; i < n_lat; i++) {
            if (done[i]) continue;
            bool lost = false;
            for (int k = 0; k < kProbe; k++) lost |= st[(int64_t)i * kProbe + k] != BMX_STATE_ALIVE;
            if (lost) {
                status[i] = 2;
                done[i] = 1;
                active--;
                continue;
            }
            // one-turn matrix by central differences: column d from probes 2d+1, 2d+2
            double *M = &jac[36 * (size_t)i];
            for (int d = 0; d < 6; d++) {
                const int64_t jp = (int64_t)i * kProbe + 2 * d + 1, jm = jp + 1;
                for (int r = 0; r < 6; r++) M[6 * r + d] = (c[r][jp] - c[r][jm]) / (2 * kStep[d]);
            }
            // Newton step on the n_dim closed coordinates (pz held for n_dim = 4)
            const int64_t j0 = (int64_t)i * kProbe;
            double a[36], b[6];
            const int nd = n_dim == 4 ? 4 : 6;
            for (int r = 0; r < nd; r++) {
                for (int k = 0; k < nd; k++) a[nd * r + k] = M[6 * r + k] - (r == k ? 1 : 0);
                b[r] = -(c[r][j0] - x[6 * i + r]);
            }
            if (!solve(nd, a, b)) {
                status[i] = 3;
                done[i] = 1;
                active--;
                continue;
            }
            // closed_orbit_calc.f90:521-534: converged when every step is below
            // |orbit| rel_tol + abs_tol
            bool small = true;
            for (int r = 0; r < nd; r++) {
                small &= std::fabs(b[r]) < std::fabs(x[6 * i + r]) * rel_tol + abs_tol;
                x[6 * i + r] += b[r];
            }
            if (small) {
                status[i] = 0;
                done[i] = 1;
                active--;
            }
        }
    }
    bmx_batch_destroy(batch);
    bmx_beam_destroy(beam);
    if (rc != BMX_OK) return rc;
    std::memcpy(orbit, x.data(), x.size() * sizeof(double));
    if (mat) std::memcpy(mat, jac.data(), jac.size() * sizeof(double));
    return BMX_OK;
}

}  // extern "C"
