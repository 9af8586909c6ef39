// Closed orbits and one-turn matrices of many lattices at once (bmad/code/closed_orbit_calc.f90:192).
//
// The reference finds the fixed point of the one-turn map with a Levenberg-Marquardt search
// (super_mrqmin) whose Jacobian is the one-turn transfer matrix built from every element's
// analytic mat6 (transfer_matrix_calc). Here the one-turn map is the GPU tracking itself: each
// lattice gets 25 particles per iteration, the trial orbit and displacements of +-h and +-2h in
// each of the six coordinates, all lattices in one batch (one captured launch sequence per
// iteration). The five-point central differences give the one-turn matrix to about 1e-9
// relative (the rounding of the tracked coordinates over the step; the truncation error, the
// step to the fourth power times the map's fifth derivatives, stays below it even on rings with
// strong sextupoles, where the three-point difference's step-squared error reached 1e-5), and a
// Newton step on
// (M - 1) d = -(f(x) - x) moves the orbit; the same convergence test as the reference ends it.
// n_dim = 4 closes x, px, y, py at the given pz with the rf cavities off (their voltage set to
// zero, as closed_orbit_calc turns them off); n_dim = 6 closes all six with the rf on.
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "bmx_internal.hpp"

using namespace bmx;

namespace {

constexpr int kProbe = 25;  // trial orbit + 6 x (+h, -h, +2h, -2h)

// Finite-difference steps: large enough that the rounding of the tracked coordinates (about
// 1e-15 after a turn) divided by the step stays near 1e-9, small enough that the five-point
// difference's truncation error does too.
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

// ---- normal-mode tunes (mode3_mod.f90: make_N's eigen_decomp_6mat and
// order_evecs_by_plane_dominance) ------------------------------------------------------------
using ld = long double;
using cld = std::complex<ld>;

// Characteristic polynomial det(lambda I - A) = sum c[k] lambda^k (c[n] = 1) of an n x n matrix,
// by the Faddeev-LeVerrier recursion in long double.
void char_poly(int n, const ld *a, ld *c)
{
    ld m[36] = {0}, t[36];
    c[n] = 1;
    for (int k = 1; k <= n; k++) {
        // m = a m + c[n-k+1] I
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                ld s = 0;
                for (int l = 0; l < n; l++) s += a[i * n + l] * m[l * n + j];
                t[i * n + j] = s + (i == j ? c[n - k + 1] : 0);
            }
        std::memcpy(m, t, sizeof(ld) * n * n);
        ld tr = 0;
        for (int i = 0; i < n; i++)
            for (int l = 0; l < n; l++) tr += a[i * n + l] * m[l * n + i];
        c[n - k] = -tr / k;
    }
}

// Roots of a monic polynomial of degree n (n <= 6) by the Aberth-Ehrlich iteration.
bool poly_roots(int n, const ld *c, cld *z)
{
    for (int i = 0; i < n; i++) z[i] = std::polar((ld)1, (ld)(2 * 3.14159265358979323846L * i / n + 0.4L));
    for (int it = 0; it < 500; it++) {
        ld worst = 0;
        for (int i = 0; i < n; i++) {
            cld p = c[n], dp = 0;
            for (int k = n - 1; k >= 0; k--) {
                dp = dp * z[i] + p;
                p = p * z[i] + c[k];
            }
            if (std::abs(dp) == 0) return false;
            const cld r = p / dp;
            cld sum = 0;
            for (int j = 0; j < n; j++)
                if (j != i) sum += (ld)1 / (z[i] - z[j]);
            const cld w = r / ((ld)1 - r * sum);
            z[i] -= w;
            worst = std::max(worst, std::abs(w));
        }
        if (worst < 1e-17L) return true;
    }
    return true;  // a last step of ~1e-17 is below what the tunes need
}

// An eigenvector of a for eigenvalue lam: Gaussian elimination of (a - lam I) with complete
// pivoting, the free variable (the last pivot's column) set to 1; then LAPACK's normalisation
// (unit Euclidean norm, the component of largest modulus real and positive).
void eigenvector(int n, const ld *a, cld lam, cld *v)
{
    cld b[36];
    int col[6];
    for (int i = 0; i < n; i++) {
        col[i] = i;
        for (int j = 0; j < n; j++) b[i * n + j] = a[i * n + j] - (i == j ? lam : (ld)0);
    }
    for (int k = 0; k < n - 1; k++) {
        int pr = k, pc = k;
        for (int i = k; i < n; i++)
            for (int j = k; j < n; j++)
                if (std::abs(b[i * n + j]) > std::abs(b[pr * n + pc])) {
                    pr = i;
                    pc = j;
                }
        for (int j = 0; j < n; j++) std::swap(b[k * n + j], b[pr * n + j]);
        for (int i = 0; i < n; i++) std::swap(b[i * n + k], b[i * n + pc]);
        std::swap(col[k], col[pc]);
        for (int i = k + 1; i < n; i++) {
            const cld f = b[i * n + k] / b[k * n + k];
            for (int j = k; j < n; j++) b[i * n + j] -= f * b[k * n + j];
        }
    }
    cld y[6];
    y[n - 1] = 1;
    for (int k = n - 2; k >= 0; k--) {
        cld s = 0;
        for (int j = k + 1; j < n; j++) s += b[k * n + j] * y[j];
        y[k] = -s / b[k * n + k];
    }
    for (int k = 0; k < n; k++) v[col[k]] = y[k];
    ld norm = 0;
    int big = 0;
    for (int k = 0; k < n; k++) {
        norm += std::norm(v[k]);
        if (std::abs(v[k]) > std::abs(v[big])) big = k;
    }
    const cld ph = std::conj(v[big]) / (std::abs(v[big]) * std::sqrt(norm));
    for (int k = 0; k < n; k++) v[k] *= ph;
}

int mode_tunes(const double *mat, int n_dim, double *tunes)
{
    const int n = n_dim, m = n_dim / 2;
    ld a[36], c[7];
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) a[i * n + j] = mat[6 * i + j];
    char_poly(n, a, c);
    cld z[6];
    if (!poly_roots(n, c, z)) return BMX_ERR_UNSTABLE;
    // one member of each conjugate pair (the one with Im > 0, as LAPACK lists it first)
    cld lam[3];
    cld vec[3][6];
    int found = 0;
    for (int i = 0; i < n; i++) {
        if (z[i].imag() <= 0) continue;
        if (std::fabs(std::abs(z[i]) - 1) > 1e-6L || z[i].imag() < 1e-12L) return BMX_ERR_UNSTABLE;
        if (found == m) return BMX_ERR_UNSTABLE;
        lam[found] = z[i];
        eigenvector(n, a, z[i], vec[found]);
        found++;
    }
    if (found != m) return BMX_ERR_UNSTABLE;
    double tune[3], re_x[3], re_y[3];
    for (int k = 0; k < m; k++) {
        // Im(v^T S conj(v)) < 0: the pair's other member (eigen_decomp_6mat, Ohmi eq. 79)
        cld q = 0;
        for (int p = 0; p < m; p++)
            q += vec[k][2 * p] * std::conj(vec[k][2 * p + 1]) - vec[k][2 * p + 1] * std::conj(vec[k][2 * p]);
        cld l = lam[k];
        if (q.imag() < 0) l = std::conj(l);
        double t = std::atan2(-(double)l.imag(), (double)l.real());
        if (t < 0) t += 2 * 3.141592653589793238462643383279;
        tune[k] = t;
        re_x[k] = std::fabs((double)vec[k][0].real());
        re_y[k] = std::fabs((double)vec[k][2].real());
    }
    int p1 = 0;
    for (int k = 1; k < m; k++)
        if (re_x[k] > re_x[p1]) p1 = k;
    int p2 = -1;
    for (int k = 0; k < m; k++)
        if (k != p1 && (p2 < 0 || re_y[k] > re_y[p2])) p2 = k;
    tunes[0] = tune[p1];
    tunes[1] = tune[p2];
    if (m == 3) tunes[2] = tune[3 - p1 - p2];
    return BMX_OK;
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
    // the tracked lattices: the rf cavities switched off for n_dim = 4; radiation
    // fluctuations off, and damping too for n_dim = 4 (closed_orbit_calc.f90:264, 322)
    Lattices held;
    std::vector<bmx_lattice *> use(lats, lats + n_lat);
    for (int32_t i = 0; i < n_lat; i++) {
        HostTable t = lats[i]->host;
        bool changed = false;
        const uint32_t rad_keep = n_dim == 4 ? ~(RAD_FLUCT | RAD_DAMP) : ~RAD_FLUCT;
        if (t.sp.rad.flags & ~rad_keep) {
            t.sp.rad.flags &= rad_keep;
            changed = true;
        }
        if (n_dim == 4) {
            for (const EleHdr &h : t.hdr)
                if (h.kind == K_RF) {
                    const int32_t q = h.off + ((h.flags & F_OFFSET) ? kOffsetLen : 0) +
                                      ((h.flags & (F_APE_ENT | F_APE_EXIT | F_APE_ENT_SAME)) ? 4 : 0) +
                                      ((h.flags & F_APE_MOMENTUM) ? 8 : 0);
                    t.prm[q + 1] = 0;  // voltage
                    changed = true;
                }
        }
        if (!changed) continue;
        bmx_lattice *L = nullptr;
        const int32_t rc = lattice_upload(ctx, std::move(t), &L);
        if (rc != BMX_OK) return rc;
        held.owned.push_back(L);
        use[i] = L;
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
                // probes 4d+1 .. 4d+4: +h, -h, +2h, -2h in coordinate d
                if (k > 0) {
                    const double f[4] = {1, -1, 2, -2};
                    v[(k - 1) / 4] += f[(k - 1) % 4] * kStep[(k - 1) / 4];
                }
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
            // one-turn matrix by five-point central differences: column d from probes 4d+1 ..
            // 4d+4, (8 (f(+h) - f(-h)) - (f(+2h) - f(-2h))) / 12 h
            double *M = &jac[36 * (size_t)i];
            for (int d = 0; d < 6; d++) {
                const int64_t j1 = (int64_t)i * kProbe + 4 * d + 1;
                for (int r = 0; r < 6; r++)
                    M[6 * r + d] = (8 * (c[r][j1] - c[r][j1 + 1]) - (c[r][j1 + 2] - c[r][j1 + 3])) /
                                   (12 * kStep[d]);
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

int32_t bmx_normal_mode_tunes(const double *mat, int32_t n_dim, double *tunes)
{
    if (!mat || !tunes || (n_dim != 4 && n_dim != 6))
        return set_error(BMX_ERR_ARG, "bmx_normal_mode_tunes: bad argument");
    for (int i = 0; i < 36; i++)
        if (!std::isfinite(mat[i])) return set_error(BMX_ERR_ARG, "bmx_normal_mode_tunes: non-finite matrix");
    const int rc = mode_tunes(mat, n_dim, tunes);
    if (rc != BMX_OK) return set_error(rc, "bmx_normal_mode_tunes: an eigenvalue is off the unit circle (unstable)");
    return BMX_OK;
}

int32_t bmx_chromaticity(bmx_ctx *ctx, bmx_lattice *const *lats, int32_t n_lat, double delta_e,
                         double pz, double *chrom, double *tune, int32_t *status)
{
    if (!ctx || n_lat < 0 || (n_lat > 0 && (!lats || !chrom || !status)) || !std::isfinite(pz) ||
        !std::isfinite(delta_e))
        return set_error(BMX_ERR_ARG, "bmx_chromaticity: bad argument");
    if (n_lat == 0) return BMX_OK;
    // chrom_calc.f90:88-93
    if (delta_e <= 0) delta_e = 1.0e-4;
    double de_low = delta_e;
    if (pz - de_low < -0.99) de_low = std::max(0.0, pz + 0.99);
    const double de = delta_e + de_low;
    // closed orbits with the rf off at the low and high energies and at pz itself
    const int32_t n3 = 3 * n_lat;
    std::vector<bmx_lattice *> all(n3);
    std::vector<double> orbit(6 * (size_t)n3, 0.0), mat(36 * (size_t)n3, 0.0);
    std::vector<int32_t> st(n3, 0);
    const double pzs[3] = {pz - de_low, pz + delta_e, pz};
    for (int32_t i = 0; i < n_lat; i++)
        for (int k = 0; k < 3; k++) {
            all[3 * i + k] = lats[i];
            orbit[6 * (3 * i + k) + 5] = pzs[k];
        }
    const int32_t rc = bmx_closed_orbit(ctx, all.data(), n3, 4, orbit.data(), mat.data(), 50,
                                        1e-12, 1e-15, st.data());
    if (rc != BMX_OK) return rc;
    for (int32_t i = 0; i < n_lat; i++) {
        status[i] = 0;
        double t[3][2];
        for (int k = 0; k < 3 && status[i] == 0; k++) {
            if (st[3 * i + k] != 0) {
                status[i] = st[3 * i + k];
                break;
            }
            if (mode_tunes(&mat[36 * (size_t)(3 * i + k)], 4, t[k]) != BMX_OK) status[i] = 4;
        }
        const double twopi = 2 * 3.141592653589793238462643383279;
        for (int m = 0; m < 2; m++) {
            double d = (t[1][m] - t[0][m]) / twopi;
            if (d > 0.5) d -= 1;
            if (d < -0.5) d += 1;
            chrom[2 * i + m] = status[i] == 0 ? d / de : NAN;
            if (tune) tune[2 * i + m] = status[i] == 0 ? t[2][m] / twopi : NAN;
        }
    }
    return BMX_OK;
}

}  // extern "C"
