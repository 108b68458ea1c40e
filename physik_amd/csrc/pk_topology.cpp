// Host-side topology builder. See pk_topology.h.
//
// Reference semantics restated here (file:line in the reference tree):
//   rest data ........... FEMModel.cpp:504-548 (1e-8 volume floor, Mat3::Inverse 1e-6 floor)
//   Lame parameters ..... FEMModel.cpp:174-200
//   lumped mass ......... FEMModel.cpp:681-713 + TetMeshPhysicsComponent.cpp:127-152
//   mass validation ..... World.cpp:560-580
//   dynamic numbering ... SolverData.cpp:239-258 (monotone over node index)
//   pattern ............. SolverData.cpp:308-405 + SparseBlockMatrix.cpp:83-129:
//                         diagonal of every dynamic node plus the canonical
//                         (min, max) pair of every two dynamic nodes sharing a
//                         tet (inactive tets included), rows sorted, unique
//   element->slot ....... TetMeshPhysicsComponent.cpp:98-124: tet t adds its
//                         block for local (a, b) with g_a <= g_b to slot
//                         (dyn a, dyn b); slots accumulate in tet order
#include "pk_topology.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace pk
{

const int kPairIndex[4][4] = {{0, 1, 2, 3}, {1, 4, 5, 6}, {2, 5, 7, 8}, {3, 6, 8, 9}};

void lame(double young, double poisson, double* lambda, double* mu)
{
    const double e = std::max(0.0, young);
    if (!(e > 0.0))
    {
        *lambda = 0.0;
        *mu = 0.0;
        return;
    }
    const double nu = std::max(static_cast<double>(-0.99f), std::min(static_cast<double>(0.49f), poisson));
    *lambda = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    *mu = e / (2.0 * (1.0 + nu));
}

namespace
{
inline void cross3(const double* a, const double* b, double* o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
inline double dot3(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
}  // namespace

void compute_tet_params(const pk_mesh_desc& d, std::vector<double>& prm)
{
    const int T = d.tet_count;
    prm.assign(static_cast<size_t>(P_COUNT) * T, 0.0);
    auto P = [&](int k, int t) -> double& { return prm[static_cast<size_t>(k) * T + t]; };
    for (int t = 0; t < T; ++t)
    {
        const int* n = d.tets + 4 * t;
        const double* p0 = d.rest + 3 * n[0];
        double c[3][3];  // columns of Dm
        for (int a = 1; a < 4; ++a)
            for (int i = 0; i < 3; ++i)
            {
                c[a - 1][i] = d.rest[3 * n[a] + i] - p0[i];
                P(P_EDGE + 3 * (a - 1) + i, t) = c[a - 1][i];
            }
        double r0[3], r1[3], r2[3];
        cross3(c[1], c[2], r0);
        cross3(c[2], c[0], r1);
        cross3(c[0], c[1], r2);
        const double det = dot3(c[0], r0);
        const double vol = std::abs(det) / 6.0;
        double lam = 0.0, mu = 0.0;
        lame(d.young[t], d.poisson[t], &lam, &mu);
        P(P_LAMBDA, t) = lam;
        P(P_MU, t) = mu;
        P(P_DAMP, t) = d.damping[t];
        if (vol <= static_cast<double>(1.0e-8f))
        {
            P(P_VOL, t) = 0.0;
            continue;  // Dm^-1 stays zero
        }
        P(P_VOL, t) = vol;
        if (std::abs(det) <= static_cast<double>(0.000001f))
        {
            for (int i = 0; i < 3; ++i)  // Mat3::Inverse floor -> Identity
                P(P_DMINV + 4 * i, t) = 1.0;
            continue;
        }
        const double id = 1.0 / det;
        for (int j = 0; j < 3; ++j)
        {
            P(P_DMINV + 0 + j, t) = r0[j] * id;
            P(P_DMINV + 3 + j, t) = r1[j] * id;
            P(P_DMINV + 6 + j, t) = r2[j] * id;
        }
    }
}

int build_topology(const pk_mesh_desc& d, Topology& topo, std::string& err, bool wantScatter)
{
    const int N = d.node_count, T = d.tet_count;
    if (N < 0 || T < 0 || (N > 0 && !d.rest) || (T > 0 && (!d.tets || !d.young || !d.poisson || !d.damping ||
                                                            !d.density)))
    {
        err = "mesh descriptor is missing arrays";
        return PK_ERR_ARG;
    }
    if (T >= (1 << 26))
    {
        err = "tet count exceeds the 2^26 limit of the packed gather lists";
        return PK_ERR_ARG;
    }
    topo = Topology();
    topo.N = N;
    topo.T = T;
    topo.tets.assign(d.tets, d.tets + 4 * static_cast<size_t>(T));
    for (int t = 0; t < T; ++t)
        for (int a = 0; a < 4; ++a)
            if (d.tets[4 * t + a] < 0 || d.tets[4 * t + a] >= N)
            {
                err = "tet " + std::to_string(t) + " references a node out of range";
                return PK_ERR_ARG;
            }

    // groups
    std::vector<int> group(static_cast<size_t>(N), 0);
    int G = 1;
    if (d.node_group && N > 0)
    {
        for (int i = 0; i < N; ++i)
        {
            group[i] = d.node_group[i];
            if (group[i] < 0 || (i > 0 && group[i] < group[i - 1]))
            {
                err = "node_group must be non-decreasing and non-negative";
                return PK_ERR_ARG;
            }
        }
        G = group[N - 1] + 1;
        for (int t = 0; t < T; ++t)
            for (int a = 1; a < 4; ++a)
                if (group[d.tets[4 * t + a]] != group[d.tets[4 * t]])
                {
                    err = "tet " + std::to_string(t) + " spans two groups";
                    return PK_ERR_ARG;
                }
    }
    topo.G = G;
    topo.nodeGroup = group;

    compute_tet_params(d, topo.prm);
    topo.poisson.assign(d.poisson, d.poisson + T);
    const double* vol = topo.prm.data() + static_cast<size_t>(P_VOL) * T;

    auto fixed = [&](int n) { return d.node_fixed && d.node_fixed[n]; };
    auto nodeActive = [&](int n) { return !d.node_active || d.node_active[n]; };
    topo.nodeActive.resize(static_cast<size_t>(N));
    topo.nodeFixed.assign(static_cast<size_t>(N), 0);
    for (int n = 0; n < N; ++n)
    {
        topo.nodeActive[n] = nodeActive(n) ? 1 : 0;
        topo.nodeFixed[n] = fixed(n) ? 1 : 0;
    }
    if (d.extra_pair_count < 0 || (d.extra_pair_count > 0 && !d.extra_pairs))
    {
        err = "mesh descriptor: bad extra pairs";
        return PK_ERR_ARG;
    }
    for (int e = 0; e < 2 * d.extra_pair_count; ++e)
        if (d.extra_pairs[e] < 0 || d.extra_pairs[e] >= N)
        {
            err = "extra pair " + std::to_string(e / 2) + " references a node out of range";
            return PK_ERR_ARG;
        }
    auto tetActive = [&](int t) { return !d.tet_active || d.tet_active[t]; };

    topo.tetActive.assign(static_cast<size_t>(T), 0);
    for (int t = 0; t < T; ++t)
        topo.tetActive[t] = tetActive(t) ? 1 : 0;
    topo.live.assign(static_cast<size_t>(T), 0);
    topo.tetShare.assign(static_cast<size_t>(T), 0.0);
    topo.mass.assign(static_cast<size_t>(N), 0.0);
    for (int t = 0; t < T; ++t)
    {
        const double rho = std::max(0.0, d.density[t]);
        if (std::isfinite(rho) && std::isfinite(vol[t]) && vol[t] > 0.0)
        {
            const double nm = rho * vol[t] * static_cast<double>(0.25f);
            if (std::isfinite(nm) && nm > 0.0)
                topo.tetShare[t] = nm;
        }
        if (!tetActive(t))
            continue;
        if (vol[t] > 0.0)
            topo.live[t] = 1;
        const double nm = topo.tetShare[t];
        if (!(nm > 0.0))
            continue;
        for (int a = 0; a < 4; ++a)
        {
            const int n = d.tets[4 * t + a];
            if (!fixed(n))
                topo.mass[n] += nm;
        }
    }
    topo.listed = topo.live;
    topo.massErrorNode = -1;
    for (int n = 0; n < N; ++n)
        if (nodeActive(n) && !fixed(n) && !(std::isfinite(topo.mass[n]) && topo.mass[n] > 0.0))
        {
            topo.massErrorNode = n;
            break;
        }
    topo.dyn.assign(static_cast<size_t>(N), -1);
    topo.nd = 0;
    for (int n = 0; n < N; ++n)
        if (nodeActive(n) && !fixed(n) && std::isfinite(topo.mass[n]) && topo.mass[n] > 0.0)
        {
            topo.dyn[n] = topo.nd++;
            topo.rowNode.push_back(n);
        }
    const int nd = topo.nd;
    topo.groupRowBegin.assign(static_cast<size_t>(G), 0);
    topo.groupRowEnd.assign(static_cast<size_t>(G), 0);
    {
        std::vector<int> cnt(static_cast<size_t>(G), 0);
        for (int r = 0; r < nd; ++r)
            ++cnt[group[topo.rowNode[r]]];
        int acc = 0;
        for (int g = 0; g < G; ++g)
        {
            topo.groupRowBegin[g] = acc;
            acc += cnt[g];
            topo.groupRowEnd[g] = acc;
        }
    }

    // Upper pattern: per-row column lists, then sort + unique.
    std::vector<int> cnt(static_cast<size_t>(nd) + 1, 0);
    for (int r = 0; r < nd; ++r)
        ++cnt[r];
    for (int t = 0; t < T; ++t)
        for (int a = 0; a < 4; ++a)
        {
            const int da = topo.dyn[d.tets[4 * t + a]];
            if (da < 0)
                continue;
            for (int b = a; b < 4; ++b)
            {
                const int db = topo.dyn[d.tets[4 * t + b]];
                if (db >= 0)
                    ++cnt[std::min(da, db)];
            }
        }
    for (int e = 0; e < d.extra_pair_count; ++e)
    {
        const int da = topo.dyn[d.extra_pairs[2 * e]], db = topo.dyn[d.extra_pairs[2 * e + 1]];
        if (da >= 0 && db >= 0)
            ++cnt[std::min(da, db)];
    }
    std::vector<long long> raw_start(static_cast<size_t>(nd) + 1, 0);
    for (int r = 0; r < nd; ++r)
        raw_start[r + 1] = raw_start[r] + cnt[r];
    std::vector<int> raw(static_cast<size_t>(raw_start[nd]));
    std::vector<long long> fillp(raw_start.begin(), raw_start.end() - 1);
    for (int r = 0; r < nd; ++r)
        raw[fillp[r]++] = r;
    for (int t = 0; t < T; ++t)
        for (int a = 0; a < 4; ++a)
        {
            const int da = topo.dyn[d.tets[4 * t + a]];
            if (da < 0)
                continue;
            for (int b = a; b < 4; ++b)
            {
                const int db = topo.dyn[d.tets[4 * t + b]];
                if (db >= 0)
                    raw[fillp[std::min(da, db)]++] = std::max(da, db);
            }
        }
    for (int e = 0; e < d.extra_pair_count; ++e)
    {
        const int da = topo.dyn[d.extra_pairs[2 * e]], db = topo.dyn[d.extra_pairs[2 * e + 1]];
        if (da >= 0 && db >= 0)
            raw[fillp[std::min(da, db)]++] = std::max(da, db);
    }
    topo.ustart.assign(static_cast<size_t>(nd) + 1, 0);
    topo.ucol.clear();
    topo.ucol.reserve(raw.size() / 2 + 1);
    for (int r = 0; r < nd; ++r)
    {
        auto b = raw.begin() + raw_start[r], e = raw.begin() + raw_start[r + 1];
        std::sort(b, e);
        e = std::unique(b, e);
        topo.ucol.insert(topo.ucol.end(), b, e);
        topo.ustart[r + 1] = static_cast<int>(topo.ucol.size());
    }
    raw.clear();
    raw.shrink_to_fit();
    const long long nnzu = topo.nnzu();
    if (nnzu >= (1LL << 31) / 9)
    {
        err = "pattern too large";
        return PK_ERR_ARG;
    }

    // Full pattern: row c = [r < c with (r, c) upper, ascending] ++ [upper row c].
    std::vector<int> lowerCount(static_cast<size_t>(nd), 0);
    for (int r = 0; r < nd; ++r)
        for (int k = topo.ustart[r]; k < topo.ustart[r + 1]; ++k)
            if (topo.ucol[k] != r)
                ++lowerCount[topo.ucol[k]];
    topo.fstart.assign(static_cast<size_t>(nd) + 1, 0);
    for (int r = 0; r < nd; ++r)
        topo.fstart[r + 1] = topo.fstart[r] + lowerCount[r] + (topo.ustart[r + 1] - topo.ustart[r]);
    topo.fcol.assign(static_cast<size_t>(topo.fstart[nd]), 0);
    topo.uFull.assign(static_cast<size_t>(nnzu), -1);
    topo.uMirror.assign(static_cast<size_t>(nnzu), -1);
    std::vector<int> lowerFill(static_cast<size_t>(nd), 0);
    for (int r = 0; r < nd; ++r)
        for (int k = topo.ustart[r]; k < topo.ustart[r + 1]; ++k)
        {
            const int c = topo.ucol[k];
            const int pos = topo.fstart[r] + lowerCount[r] + (k - topo.ustart[r]);
            topo.fcol[pos] = c;
            topo.uFull[k] = pos;
            if (c != r)
            {
                const int mpos = topo.fstart[c] + lowerFill[c]++;
                topo.fcol[mpos] = r;
                topo.uMirror[k] = mpos;
            }
        }

    auto slot_of = [&](int r, int c) -> int {
        if (r > c)
            std::swap(r, c);
        const int* b = topo.ucol.data() + topo.ustart[r];
        const int* e = topo.ucol.data() + topo.ustart[r + 1];
        const int* it = std::lower_bound(b, e, c);
        return (it != e && *it == c) ? static_cast<int>(it - topo.ucol.data()) : -1;
    };

    // Gather lists: upper slot <- (tet, pair, transpose) over live tets in order.
    std::vector<int> slots10(static_cast<size_t>(T) * 10, -1);
    std::vector<int> ccount(static_cast<size_t>(nnzu) + 1, 0);
    for (int t = 0; t < T; ++t)
    {
        if (!topo.live[t])
            continue;
        for (int a = 0; a < 4; ++a)
            for (int b = a; b < 4; ++b)
            {
                const int da = topo.dyn[d.tets[4 * t + a]], db = topo.dyn[d.tets[4 * t + b]];
                if (da < 0 || db < 0)
                    continue;
                const int s = slot_of(da, db);
                slots10[10 * static_cast<size_t>(t) + kPairIndex[a][b]] = s;
                ++ccount[s];
            }
    }
    topo.cstart.assign(static_cast<size_t>(nnzu) + 1, 0);
    for (long long s = 0; s < nnzu; ++s)
        topo.cstart[s + 1] = topo.cstart[s] + ccount[s];
    topo.clist.assign(static_cast<size_t>(topo.cstart[nnzu]), 0);
    std::vector<int> cfill(topo.cstart.begin(), topo.cstart.end() - 1);
    for (int t = 0; t < T; ++t)
    {
        if (!topo.live[t])
            continue;
        for (int a = 0; a < 4; ++a)
            for (int b = a; b < 4; ++b)
            {
                const int s = slots10[10 * static_cast<size_t>(t) + kPairIndex[a][b]];
                if (s < 0)
                    continue;
                const int da = topo.dyn[d.tets[4 * t + a]], db = topo.dyn[d.tets[4 * t + b]];
                // the slot's row node is the smaller dynamic id; the stored
                // block is oriented (a, b), so flip when b owns the row
                topo.clist[cfill[s]++] = pack_contrib(t, kPairIndex[a][b], da > db ? 1 : 0);
            }
    }


    // rhs gather lists: dynamic row <- (tet, corner) over live tets in order.
    std::vector<int> ncount(static_cast<size_t>(nd) + 1, 0);
    for (int t = 0; t < T; ++t)
        if (topo.live[t])
            for (int a = 0; a < 4; ++a)
            {
                const int da = topo.dyn[d.tets[4 * t + a]];
                if (da >= 0)
                    ++ncount[da];
            }
    topo.nstart.assign(static_cast<size_t>(nd) + 1, 0);
    for (int r = 0; r < nd; ++r)
        topo.nstart[r + 1] = topo.nstart[r] + ncount[r];
    topo.nlist.assign(static_cast<size_t>(topo.nstart[nd]), 0);
    std::vector<int> nfill(topo.nstart.begin(), topo.nstart.end() - 1);
    for (int t = 0; t < T; ++t)
        if (topo.live[t])
            for (int a = 0; a < 4; ++a)
            {
                const int da = topo.dyn[d.tets[4 * t + a]];
                if (da >= 0)
                    topo.nlist[nfill[da]++] = pack_corner(t, a);
            }
    topo.cpos.assign(static_cast<size_t>(T) * 4, -1);
    for (size_t k = 0; k < topo.nlist.size(); ++k)
        topo.cpos[static_cast<size_t>(topo.nlist[k] & 3) * T + (topo.nlist[k] >> 2)] = static_cast<int>(k);

    // reverse-mode gather lists: node <- (tet, corner) over live tets in order
    topo.astart.assign(static_cast<size_t>(N) + 1, 0);
    for (int t = 0; t < T; ++t)
        if (topo.live[t])
            for (int a = 0; a < 4; ++a)
                ++topo.astart[d.tets[4 * t + a] + 1];
    for (int n = 0; n < N; ++n)
        topo.astart[n + 1] += topo.astart[n];
    topo.alist.assign(static_cast<size_t>(topo.astart[N]), 0);
    {
        std::vector<int> afill(topo.astart.begin(), topo.astart.end() - 1);
        for (int t = 0; t < T; ++t)
            if (topo.live[t])
                for (int a = 0; a < 4; ++a)
                    topo.alist[afill[d.tets[4 * t + a]]++] = pack_corner(t, a);
    }
    topo.apos.assign(static_cast<size_t>(T) * 4, -1);
    for (size_t k = 0; k < topo.alist.size(); ++k)
        topo.apos[static_cast<size_t>(topo.alist[k] & 3) * T + (topo.alist[k] >> 2)] = static_cast<int>(k);

    if (wantScatter)
    {
        topo.scatter.assign(static_cast<size_t>(T) * 16, -1);
        for (int t = 0; t < T; ++t)
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                {
                    const int da = topo.dyn[d.tets[4 * t + a]], db = topo.dyn[d.tets[4 * t + b]];
                    topo.scatter[16 * static_cast<size_t>(t) + 4 * a + b] = (da < 0 || db < 0) ? -1 : slot_of(da, db);
                }
    }
    return PK_OK;
}

}  // namespace pk

// ---- host-only C ABI ----

struct pk_pattern
{
    pk::Topology topo;
};

namespace pk
{
void set_error(const std::string& msg);
}

extern "C" int pk_pattern_create(const pk_mesh_desc* desc, pk_pattern** out)
{
    if (!desc || !out)
    {
        pk::set_error("null argument");
        return PK_ERR_ARG;
    }
    auto* p = new pk_pattern();
    std::string err;
    const int rc = pk::build_topology(*desc, p->topo, err, true);
    if (rc != PK_OK)
    {
        pk::set_error(err);
        delete p;
        *out = nullptr;
        return rc;
    }
    *out = p;
    return PK_OK;
}

extern "C" void pk_pattern_destroy(pk_pattern* p) { delete p; }

extern "C" int pk_pattern_sizes(const pk_pattern* p, int* nd, long long* nnzu, long long* nnzf, long long* contrib)
{
    if (!p)
        return PK_ERR_ARG;
    if (nd) *nd = p->topo.nd;
    if (nnzu) *nnzu = p->topo.nnzu();
    if (nnzf) *nnzf = p->topo.nnzf();
    if (contrib) *contrib = static_cast<long long>(p->topo.clist.size());
    return PK_OK;
}

extern "C" int pk_pattern_copy(const pk_pattern* p, int* dyn, int* upper_start, int* upper_col, int* full_start,
                               int* full_col, int* scatter)
{
    if (!p)
        return PK_ERR_ARG;
    const pk::Topology& t = p->topo;
    auto cp = [](int* dst, const std::vector<int>& src) {
        if (dst && !src.empty())
            std::memcpy(dst, src.data(), src.size() * sizeof(int));
    };
    cp(dyn, t.dyn);
    cp(upper_start, t.ustart);
    cp(upper_col, t.ucol);
    cp(full_start, t.fstart);
    cp(full_col, t.fcol);
    cp(scatter, t.scatter);
    return PK_OK;
}
