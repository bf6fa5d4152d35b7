// Host-side data of one solve: lead channels, transverse modes, stencil.
#pragma once

#include <map>
#include <utility>
#include <vector>

#include "rgf_internal.h"

namespace rgf {

struct Side {                       // setup_longitudinal (setup_channels.m:179-238)
    std::vector<cd> kl;             // longitudinal kxdx (2D) / kzdx (3D), all N channels
    std::vector<int32_t> ind_prop;  // 0-based propagating channel indices
    std::vector<cd> sqrt_mu;        // sqrt(sin kl) of the propagating channels
    std::vector<int32_t> conj;      // 0-based ind_prop_conj (2D, when defined)
    bool has_conj = false;
    int32_t n_prop = 0;
};

// Transverse grid: n1 (fast, 2D: y; 3D: x) by n2 (3D: y; 2D: 1).  Channel a
// is (a1, a2) = (a % n1, a / n1); its mode is f1[.][a1] * f2[.][a2].
struct Channels {
    int ndim = 2;
    int64_t n1 = 0, n2 = 1, N = 0, L_slices = 0;
    int bc[2] = {0, 0};
    double ka[2] = {0, 0};
    std::vector<double> kt1, kt2;   // transverse wavenumbers per channel
    std::vector<cd> f1, f2;         // 1D mode tables, row-major [m][a]
    cd k2, epsL, epsR;
    bool same_sides = false;
    Side L, R;
};

struct HostCSR {
    int32_t n = 0;
    std::vector<int32_t> rowptr, col;
    std::vector<double> vre, vim;
};

struct Selection {                  // 0-based channel indices, user order
    std::vector<int32_t> inL, inR, outL, outR;
};

cd kx_from_z(cd z);
double mode_norm(int64_t n, int bc);
cd mode_shape(int bc, int64_t m, double k);
void validate_problem(const rgf_problem* p);
// modes = false leaves the transverse mode tables f1, f2 empty (the channel
// counts, wavenumbers and lead dispersion do not need them); build_modes
// fills them later.
Channels setup_channels(const rgf_problem* p, bool modes = true);
void build_modes(Channels& ch);
std::map<std::pair<int64_t, int64_t>, cd> laplacian_1d(int64_t n, int bc, double ka);
HostCSR slice_stencil(const Channels& ch);
Selection make_selection(const rgf_selection* s, const Channels& ch);

}  // namespace rgf
