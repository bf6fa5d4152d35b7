// Host-side (FP64) lead channels, transverse modes and slice stencil.
//
// Restates setup_channels.m, setup_channels_3D.m and build_laplacian_1d.m of
// chiaweihsu/RGF in C++; see rgf_host.h for the data layout.
#include "rgf_host.h"

#include <algorithm>
#include <cmath>
#include <map>

namespace rgf {

static const double PI = 3.14159265358979323846;

// kxdx = 2 asin(sqrt(z)) on the physical branch (setup_channels.m:185-215).
// The reference flips the sign where MATLAB's real-argument asin(x>1) has a
// negative imaginary part; C99 casin has the opposite sign on that cut, so
// the documented rule (setup_channels.m:52-60) is applied by cases instead:
// real z in [0,1]: real kx; real z < 0: kx = 2i asinh(sqrt(-z)); real z > 1:
// kx = -pi + 2i acosh(sqrt(z)) (MATLAB's representative); complex z: principal
// value, negated where Im z < 0 and (Re z < 0 or Re z > 1).
cd kx_from_z(cd z) {
    const double x = z.real(), y = z.imag();
    if (y == 0.0) {
        if (x >= 0.0 && x <= 1.0) return cd(2.0 * std::asin(std::sqrt(x)), 0.0);
        if (x < 0.0) return cd(0.0, 2.0 * std::asinh(std::sqrt(-x)));
        return cd(-PI, 2.0 * std::acosh(std::sqrt(x)));
    }
    cd kx = 2.0 * std::asin(std::sqrt(z));
    if (y < 0.0 && (x < 0.0 || x > 1.0)) kx = -kx;
    return kx;
}

// One transverse direction: the wavenumbers (setup_channels.m:122-156).
static void transverse_k(int64_t n, int bc, double ka, std::vector<double>& k) {
    k.assign(n, 0.0);
    if (bc == RGF_BC_BLOCH) {
        const int64_t z = (n % 2 == 1) ? (n + 1) / 2 : n / 2;  // ind_zero_ky (:125-129)
        for (int64_t a = 0; a < n; ++a) k[a] = ka / n + (double)(a + 1 - z) * (2 * PI / n);
    } else if (bc == RGF_BC_DIRICHLET) {
        for (int64_t a = 0; a < n; ++a) k[a] = (double)(a + 1) * (PI / (n + 1));
    } else if (bc == RGF_BC_NEUMANN) {
        for (int64_t a = 0; a < n; ++a) k[a] = (double)a * (PI / n);
    } else {  // DN / ND
        for (int64_t a = 0; a < n; ++a) k[a] = (a + 0.5) * (PI / (n + 0.5));
    }
}

// phi(m) = mode_norm(n, bc) * mode_shape(bc, m, k): the mode with
// wavenumber k at site m (0-based) of one direction of n sites
// (channels.fun_phi, setup_channels.m:121-156); Neumann's k = 0 mode is the
// constant 1/sqrt(n).
double mode_norm(int64_t n, int bc) {
    if (bc == RGF_BC_BLOCH) return 1.0 / std::sqrt((double)n);
    if (bc == RGF_BC_DIRICHLET) return std::sqrt(2.0 / (n + 1));
    if (bc == RGF_BC_NEUMANN) return std::sqrt(2.0 / n);
    return std::sqrt(2.0 / (n + 0.5));  // Dirichlet-Neumann, Neumann-Dirichlet
}

cd mode_shape(int bc, int64_t m, double k) {
    if (bc == RGF_BC_BLOCH) return std::polar(1.0, (double)(m + 1) * k);
    if (bc == RGF_BC_DIRICHLET || bc == RGF_BC_DIRICHLET_NEUMANN) return std::sin((double)(m + 1) * k);
    if (bc == RGF_BC_NEUMANN) return std::cos((m + 0.5) * k) - (k == 0 ? (1 - std::sqrt(0.5)) : 0.0);
    return std::cos((m + 0.5) * k);  // Neumann-Dirichlet
}

// The n x n mode table f[m][a] = phi_a(m) (row-major) of one direction.
static void transverse_modes(int64_t n, int bc, const std::vector<double>& k, std::vector<cd>& f) {
    f.assign((size_t)n * n, cd(0, 0));
    const double s = mode_norm(n, bc);
    for (int64_t m = 0; m < n; ++m)
        for (int64_t a = 0; a < n; ++a) f[m * n + a] = mode_shape(bc, m, k[a]) * s;
}

void build_modes(Channels& ch) {
    std::vector<double> k;
    transverse_k(ch.n1, ch.bc[0], ch.ka[0], k);
    transverse_modes(ch.n1, ch.bc[0], k, ch.f1);
    if (ch.ndim == 3) {
        transverse_k(ch.n2, ch.bc[1], ch.ka[1], k);
        transverse_modes(ch.n2, ch.bc[1], k, ch.f2);
    } else {
        ch.f2.assign(1, cd(1, 0));
    }
}

static void longitudinal(cd k0dx2_eps, const std::vector<double>& sin2, Side& side) {
    const size_t N = sin2.size();
    side.kl.resize(N);
    side.ind_prop.clear();
    for (size_t a = 0; a < N; ++a) {
        const cd z = 0.25 * k0dx2_eps - sin2[a];
        side.kl[a] = kx_from_z(z);
        if (z.real() > 0 && z.real() < 1) side.ind_prop.push_back((int32_t)a);  // (:195)
    }
    side.n_prop = (int32_t)side.ind_prop.size();
    side.sqrt_mu.resize(side.n_prop);
    for (int32_t p = 0; p < side.n_prop; ++p) side.sqrt_mu[p] = std::sqrt(std::sin(side.kl[side.ind_prop[p]]));
}

static void conj_perm(const Channels& ch, Side& side) {
    // setup_channels.m:221-236 (2D only; undefined for Bloch ka != 0)
    const int32_t n = side.n_prop;
    side.has_conj = false;
    side.conj.clear();
    if (ch.ndim != 2) return;
    if (ch.bc[0] != RGF_BC_BLOCH) {
        for (int32_t a = 0; a < n; ++a) side.conj.push_back(a);
        side.has_conj = true;
    } else if (ch.ka[0] == 0.0) {
        const int64_t ny = ch.n1;
        const int32_t zk = (int32_t)((ny % 2 == 1) ? (ny + 1) / 2 : ny / 2) - 1;
        const bool has_zero = std::find(side.ind_prop.begin(), side.ind_prop.end(), zk) != side.ind_prop.end();
        if (has_zero || n % 2 == 0) {
            for (int32_t a = n - 1; a >= 0; --a) side.conj.push_back(a);
        } else {
            for (int32_t a = n - 2; a >= 0; --a) side.conj.push_back(a);
            side.conj.push_back(n - 1);
        }
        side.has_conj = true;
    }
}

void validate_problem(const rgf_problem* p) {
    if (!p) throw Error(RGF_ERR_INVALID_ARGUMENT, "syst must be a structure");
    if (p->ndim != 2 && p->ndim != 3) throw Error(RGF_ERR_INVALID_ARGUMENT, "ndim must be 2 or 3");
    for (int d = 0; d < 3; ++d)
        if (p->dims[d] < 0) throw Error(RGF_ERR_INVALID_ARGUMENT, "dimensions must be non-negative");
    if (p->ndim == 2 && p->dims[0] == 0)
        throw Error(RGF_ERR_INVALID_ARGUMENT, "need at least one site along the transverse (y) direction");
    if (p->ndim == 3 && p->dims[0] * p->dims[1] == 0)
        throw Error(RGF_ERR_INVALID_ARGUMENT, "need at least one site along the transverse directions (x,y)");
    const int nbc = p->ndim == 2 ? 1 : 2;
    for (int d = 0; d < nbc; ++d) {
        const int b = p->bc[d];
        if (b < RGF_BC_BLOCH || b > RGF_BC_NEUMANN_DIRICHLET)
            throw Error(RGF_ERR_BOUNDARY_CONDITION, "Boundary condition BC is not an eligible option");
        if (p->ndim == 3 && b != RGF_BC_BLOCH && b != RGF_BC_DIRICHLET)
            throw Error(RGF_ERR_BOUNDARY_CONDITION, "Boundary condition BC is not an eligible option");
        if (!std::isfinite(p->bloch_ka[d]))
            throw Error(RGF_ERR_BOUNDARY_CONDITION, "Bloch phase k*a must be a finite real number");
    }
    if (p->ndim == 3 && p->bc[0] != p->bc[1])
        throw Error(RGF_ERR_BOUNDARY_CONDITION,
                    "xyBC must be 'periodic', 'Dirichlet' or a real vector [kxa, kya]");
    const double* sc[3] = {p->eps_L, p->eps_R, p->k0dx};
    for (auto v : sc)
        if (!std::isfinite(v[0]) || !std::isfinite(v[1]))
            throw Error(RGF_ERR_INVALID_ARGUMENT, "epsilon_L, epsilon_R and k0dx must be finite numeric scalars");
    const int64_t total = p->dims[0] * p->dims[1] * (p->ndim == 3 ? p->dims[2] : 1);
    if (total > 0 && !p->eps_re) throw Error(RGF_ERR_INVALID_ARGUMENT, "syst.epsilon must be provided");
}

Channels setup_channels(const rgf_problem* p, bool modes) {
    validate_problem(p);
    Channels ch;
    ch.ndim = p->ndim;
    if (p->ndim == 2) {
        ch.n1 = p->dims[0];
        ch.n2 = 1;
        ch.L_slices = p->dims[1];
        ch.bc[0] = p->bc[0];
        ch.bc[1] = RGF_BC_DIRICHLET;
        ch.ka[0] = p->bc[0] == RGF_BC_BLOCH ? p->bloch_ka[0] : 0.0;
        ch.ka[1] = 0.0;
    } else {
        ch.n1 = p->dims[0];
        ch.n2 = p->dims[1];
        ch.L_slices = p->dims[2];
        for (int d = 0; d < 2; ++d) {
            ch.bc[d] = p->bc[d];
            ch.ka[d] = p->bc[d] == RGF_BC_BLOCH ? p->bloch_ka[d] : 0.0;
        }
    }
    ch.N = ch.n1 * ch.n2;
    std::vector<double> k1, k2;
    transverse_k(ch.n1, ch.bc[0], ch.ka[0], k1);
    if (ch.ndim == 3)
        transverse_k(ch.n2, ch.bc[1], ch.ka[1], k2);
    else
        k2.assign(1, 0.0);
    if (modes) build_modes(ch);
    ch.kt1.resize(ch.N);
    ch.kt2.resize(ch.N);
    std::vector<double> sin2(ch.N);
    for (int64_t a2 = 0; a2 < ch.n2; ++a2)
        for (int64_t a1 = 0; a1 < ch.n1; ++a1) {
            const int64_t a = a1 + ch.n1 * a2;  // column-major, x fast (setup_channels_3D.m:114-118)
            ch.kt1[a] = k1[a1];
            ch.kt2[a] = ch.ndim == 3 ? k2[a2] : 0.0;
            const double s1 = std::sin(k1[a1] / 2), s2 = ch.ndim == 3 ? std::sin(k2[a2] / 2) : 0.0;
            sin2[a] = s1 * s1 + s2 * s2;
        }
    const cd k0dx(p->k0dx[0], p->k0dx[1]);
    ch.k2 = k0dx * k0dx;
    ch.epsL = cd(p->eps_L[0], p->eps_L[1]);
    ch.epsR = cd(p->eps_R[0], p->eps_R[1]);
    longitudinal(ch.k2 * ch.epsL, sin2, ch.L);
    conj_perm(ch, ch.L);
    ch.same_sides = (ch.epsL == ch.epsR);
    if (ch.same_sides) {
        ch.R = ch.L;
    } else {
        longitudinal(ch.k2 * ch.epsR, sin2, ch.R);
        conj_perm(ch, ch.R);
    }
    return ch;
}

// build_laplacian_1d.m:46-75 as the sparse product -grad1^H grad1; entries
// with equal (row, col) accumulate, as MATLAB's sparse() does (this fixes the
// Bloch n = 1 case to -4 sin^2(ka/2)).
std::map<std::pair<int64_t, int64_t>, cd> laplacian_1d(int64_t n, int bc, double ka) {
    // rows of grad1 as lists of (col, value)
    std::vector<std::vector<std::pair<int64_t, cd>>> g;
    if (bc == RGF_BC_BLOCH) {
        g.resize(n);
        for (int64_t m = 0; m < n; ++m) {
            g[m].push_back({m, cd(1, 0)});
            if (m >= 1) g[m].push_back({m - 1, cd(-1, 0)});
        }
        if (n >= 1) g[0].push_back({n - 1, -std::exp(cd(0, -ka))});
    } else if (bc == RGF_BC_DIRICHLET) {
        g.resize(n + 1);
        for (int64_t m = 0; m <= n; ++m) {
            if (m < n) g[m].push_back({m, cd(1, 0)});
            if (m >= 1) g[m].push_back({m - 1, cd(-1, 0)});
        }
    } else if (bc == RGF_BC_NEUMANN) {
        g.resize(std::max<int64_t>(n - 1, 0));
        for (int64_t m = 0; m + 1 < n; ++m) {
            g[m].push_back({m + 1, cd(1, 0)});
            g[m].push_back({m, cd(-1, 0)});
        }
    } else if (bc == RGF_BC_DIRICHLET_NEUMANN) {
        g.resize(n);
        for (int64_t m = 0; m < n; ++m) {
            g[m].push_back({m, cd(1, 0)});
            if (m >= 1) g[m].push_back({m - 1, cd(-1, 0)});
        }
    } else {
        g.resize(n);
        for (int64_t m = 0; m < n; ++m) {
            g[m].push_back({m, cd(-1, 0)});
            if (m + 1 < n) g[m].push_back({m + 1, cd(1, 0)});
        }
    }
    std::map<std::pair<int64_t, int64_t>, cd> A;
    for (auto& row : g)
        for (auto& x : row)
            for (auto& y : row) A[{x.first, y.first}] -= std::conj(x.second) * y.second;
    return A;
}

// A0 = 2I - kron(I, L1) - kron(L2, I) (cal_smatrix_RGF.m:209; _3D.m:218-224).
// Each row holds a handful of entries: they are gathered per row in the
// order the terms are added (the diagonal, then L1, then L2, as sparse()
// sums them) and sorted by column at the end.
HostCSR slice_stencil(const Channels& ch) {
    const int64_t n1 = ch.n1, n2 = ch.n2, N = ch.N;
    std::vector<std::vector<std::pair<int64_t, cd>>> rows(N);
    auto add = [&](int64_t r, int64_t col, cd v) {
        for (auto& e : rows[r])
            if (e.first == col) {
                e.second += v;
                return;
            }
        rows[r].push_back({col, cd(0, 0) + v});
    };
    for (int64_t i = 0; i < N; ++i) add(i, i, cd(2, 0));
    auto L1 = laplacian_1d(n1, ch.bc[0], ch.ka[0]);
    for (int64_t i2 = 0; i2 < n2; ++i2)
        for (auto& e : L1) add(e.first.first + n1 * i2, e.first.second + n1 * i2, -e.second);
    if (ch.ndim == 3) {
        auto L2 = laplacian_1d(n2, ch.bc[1], ch.ka[1]);
        for (int64_t i1 = 0; i1 < n1; ++i1)
            for (auto& e : L2) add(i1 + n1 * e.first.first, i1 + n1 * e.first.second, -e.second);
    }
    HostCSR c;
    c.n = (int32_t)N;
    c.rowptr.assign(N + 1, 0);
    for (int64_t i = 0; i < N; ++i) {
        std::sort(rows[i].begin(), rows[i].end(),
                  [](const std::pair<int64_t, cd>& a, const std::pair<int64_t, cd>& b) { return a.first < b.first; });
        c.rowptr[i + 1] = c.rowptr[i] + (int32_t)rows[i].size();
    }
    c.col.reserve(c.rowptr[N]);
    c.vre.reserve(c.rowptr[N]);
    c.vim.reserve(c.rowptr[N]);
    for (auto& r : rows)
        for (auto& e : r) {
            c.col.push_back((int32_t)e.first);
            c.vre.push_back(e.second.real());
            c.vim.push_back(e.second.imag());
        }
    return c;
}

// cal_smatrix_RGF.m:103-175: counts of RGF_ALL expand to 1..N_prop; indices
// are validated against N_prop and converted to 0-based.
static std::vector<int32_t> pick(const int32_t* ind, int32_t n, int32_t nprop, const char* what, const char* side) {
    std::vector<int32_t> v;
    if (n == RGF_ALL) {
        for (int32_t a = 0; a < nprop; ++a) v.push_back(a);
        return v;
    }
    if (n < 0) throw Error(RGF_ERR_INVALID_ARGUMENT, "negative channel count");
    if (n > 0 && !ind) throw Error(RGF_ERR_INVALID_ARGUMENT, "channel index array is NULL");
    for (int32_t q = 0; q < n; ++q) {
        if (ind[q] < 1 || ind[q] > nprop) {
            char buf[200];
            snprintf(buf, sizeof buf,
                     "%s, when specified, must be an array of positive integers not exceeding N_prop_%s=%d", what,
                     side, nprop);
            throw Error(RGF_ERR_CHANNEL_INDEX, buf);
        }
        v.push_back(ind[q] - 1);
    }
    return v;
}

Selection make_selection(const rgf_selection* s, const Channels& ch) {
    if (!s) throw Error(RGF_ERR_INVALID_ARGUMENT, "selection must be provided");
    Selection r;
    r.inL = pick(s->ind_in_L, s->n_in_L, ch.L.n_prop, "in.ind_in_L", "L");
    r.inR = pick(s->ind_in_R, s->n_in_R, ch.R.n_prop, "in.ind_in_R", "R");
    r.outL = pick(s->ind_out_L, s->n_out_L, ch.L.n_prop, "out.ind_out_L", "L");
    r.outR = pick(s->ind_out_R, s->n_out_R, ch.R.n_prop, "out.ind_out_R", "R");
    return r;
}

}  // namespace rgf
