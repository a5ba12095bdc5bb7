// msx_specimen.hip -- host-side specimen generator for the uneven-job
// workload (amorphous carbon): random sequential addition in a periodic box.
//
// Candidates are uniform in the box; a candidate is accepted when no accepted
// atom lies closer than min_dist (minimum image), and the process stops at the
// target count or after max_candidates draws.  Candidate i draws its three
// coordinates from Philox-4x32-10 with counter (i lo, i hi, 0x5eed, 0) and key
// (seed lo, seed hi), so a (seed, box) pair always yields the same specimen.
// Accepted atoms are kept in a cell grid (cells >= min_dist wide), so each
// candidate tests the atoms of 27 cells: O(n) expected time.
#include <cmath>
#include <cstdint>
#include <vector>

#include "msx_internal.h"
#include "msx_philox.h"
#include "../../include/msx.h"

using msx::fail;

namespace {

struct Grid {
    int n[3];
    double cell[3];
    std::vector<std::vector<int64_t>> bins;
    Grid(const double box[3], double d) {
        for (int k = 0; k < 3; k++) {
            n[k] = (int)std::floor(box[k] / d);
            if (n[k] < 1) n[k] = 1;
            cell[k] = box[k] / n[k];
        }
        bins.resize((size_t)n[0] * n[1] * n[2]);
    }
    int idx(int i, int j, int k) const { return (i * n[1] + j) * n[2] + k; }
};

}  // namespace

extern "C" MSX_API int msx_amorphous(double a, double b, double c, int32_t Z, int64_t natoms, double min_dist,
                                     uint64_t seed, int64_t max_candidates, msx_atom* out, int64_t* nplaced) {
    if (!(a > 0) || !(b > 0) || !(c > 0)) return fail(MSX_EINVAL, "box sides must be > 0");
    if (Z < 1 || Z > 103) return fail(MSX_EINVAL, "Z = %d outside 1..103", Z);
    if (natoms < 0 || !(min_dist >= 0) || max_candidates < 0) return fail(MSX_EINVAL, "bad counts or distance");
    if (natoms > 0 && !out) return fail(MSX_EINVAL, "out is NULL");
    if (!nplaced) return fail(MSX_EINVAL, "nplaced is NULL");
    const double box[3] = {a, b, c};
    // with fewer than 3 cells along an axis the 27-cell stencil would visit
    // a cell twice: fall back to one cell along that axis (all atoms tested)
    Grid g(box, min_dist > 0 ? min_dist : std::max(a, std::max(b, c)));
    for (int k = 0; k < 3; k++)
        if (g.n[k] < 3) {
            g.n[k] = 1;
            g.cell[k] = box[k];
        }
    g.bins.assign((size_t)g.n[0] * g.n[1] * g.n[2], {});
    const double d2 = min_dist * min_dist;
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    int64_t placed = 0;
    for (int64_t i = 0; i < max_candidates && placed < natoms; i++) {
        const msx::Philox4 r = msx::philox4x32_10((uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0x5eedu, 0u, k0, k1);
        const msx::Philox4 r2 = msx::philox4x32_10((uint32_t)i, (uint32_t)((uint64_t)i >> 32), 0x5eedu, 1u, k0, k1);
        double p[3] = {msx::uniform53(r.v[0], r.v[1]) * a, msx::uniform53(r.v[2], r.v[3]) * b,
                       msx::uniform53(r2.v[0], r2.v[1]) * c};
        int ci[3];
        for (int k = 0; k < 3; k++) {
            if (p[k] >= box[k]) p[k] -= box[k];
            ci[k] = std::min(g.n[k] - 1, (int)(p[k] / g.cell[k]));
        }
        bool ok = true;
        const int sx = g.n[0] > 1 ? 1 : 0, sy = g.n[1] > 1 ? 1 : 0, sz = g.n[2] > 1 ? 1 : 0;
        for (int dx = -sx; dx <= sx && ok; dx++)
            for (int dy = -sy; dy <= sy && ok; dy++)
                for (int dz = -sz; dz <= sz && ok; dz++) {
                    const int ii = (ci[0] + dx + g.n[0]) % g.n[0], jj = (ci[1] + dy + g.n[1]) % g.n[1],
                              kk = (ci[2] + dz + g.n[2]) % g.n[2];
                    for (int64_t q : g.bins[g.idx(ii, jj, kk)]) {
                        double dd = 0;
                        const double qp[3] = {out[q].x, out[q].y, out[q].z};
                        for (int k = 0; k < 3; k++) {
                            double t = p[k] - qp[k];
                            t -= box[k] * std::nearbyint(t / box[k]);
                            dd += t * t;
                        }
                        if (dd < d2) {
                            ok = false;
                            break;
                        }
                    }
                }
        if (!ok) continue;
        out[placed].Z = Z;
        out[placed].occ = 1.0f;
        out[placed].x = p[0];
        out[placed].y = p[1];
        out[placed].z = p[2];
        g.bins[g.idx(ci[0], ci[1], ci[2])].push_back(placed);
        placed++;
    }
    *nplaced = placed;
    return MSX_OK;
}
