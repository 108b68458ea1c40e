// PHYSIK_* drop-in host API over the GPU system (physik_amd.h).
//
// World bookkeeping (nodes, components, handles, settings) lives on the host,
// as in the reference (src/Core/World/World.cpp, src/API/PhysiKAPI.cpp); the
// implicit step itself runs on the GPU. The device copy of the system is
// rebuilt when topology, flags or materials change, and node state moves
// between host and device only when one side has changed it.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <set>
#include <vector>

#include "PhysiKAPI.h"
#include "physik_amd.h"

namespace physik_amd
{

thread_local std::string g_apiError;

// Material defaults of include/PhysiK/PhysicsData/Material.h:5-11.
struct Material
{
    float density = 1.0f;
    float youngModulus = 25.0f;
    float poissonRatio = 0.3f;
    float damping = 0.25f;
};

// TetMeshGenerator output (src/Geometry/TetMeshGenerator.cpp:153-282).
struct GeneratedMesh
{
    std::vector<double> positions;  // welded nodes, xyz
    std::vector<int> tets;          // 4 per tet, local ids
};

std::unordered_set<GeneratedMesh*> g_meshes;  // PhysiKAPI.cpp:24

// Node welding within 1e-5 on a hash grid of that cell size, then removal of
// tets with invalid, duplicate or nearly coplanar nodes ((6V)^2 <= 1e-24).
// S is the arithmetic type of the checks: float for the reference ABI,
// double for the FP64 variant.
template <class S>
GeneratedMesh* generate(const S* pos, int nodeCount, const int* tets, int tetCount)
{
    auto* m = new GeneratedMesh();
    if (pos == nullptr || nodeCount <= 0)
        return m;
    const S tol = static_cast<S>(1.0e-5f);
    const S tol2 = tol * tol;
    struct Key
    {
        int x, y, z;
        bool operator==(const Key& o) const { return x == o.x && y == o.y && z == o.z; }
    };
    struct KeyHash
    {
        size_t operator()(const Key& k) const
        {
            return std::hash<int>()(k.x) ^ (std::hash<int>()(k.y) << 1u) ^ (std::hash<int>()(k.z) << 2u);
        }
    };
    std::unordered_map<Key, std::vector<int>, KeyHash> grid;
    std::vector<S> kept;
    std::vector<int> remap(static_cast<size_t>(nodeCount), -1);
    auto cell = [&](const S* p) {
        return Key{static_cast<int>(std::floor(p[0] / tol)), static_cast<int>(std::floor(p[1] / tol)),
                   static_cast<int>(std::floor(p[2] / tol))};
    };
    for (int i = 0; i < nodeCount; ++i)
    {
        const S* p = pos + 3 * i;
        const Key c = cell(p);
        int found = -1;
        for (int dz = -1; dz <= 1 && found < 0; ++dz)
            for (int dy = -1; dy <= 1 && found < 0; ++dy)
                for (int dx = -1; dx <= 1 && found < 0; ++dx)
                {
                    auto it = grid.find(Key{c.x + dx, c.y + dy, c.z + dz});
                    if (it == grid.end())
                        continue;
                    for (int nn : it->second)
                    {
                        const S ex = kept[3 * nn] - p[0], ey = kept[3 * nn + 1] - p[1], ez = kept[3 * nn + 2] - p[2];
                        if (ex * ex + ey * ey + ez * ez <= tol2)
                        {
                            found = nn;
                            break;
                        }
                    }
                }
        if (found >= 0)
        {
            remap[i] = found;
            continue;
        }
        const int nn = static_cast<int>(kept.size() / 3);
        kept.push_back(p[0]);
        kept.push_back(p[1]);
        kept.push_back(p[2]);
        remap[i] = nn;
        grid[c].push_back(nn);
    }
    m->positions.assign(kept.begin(), kept.end());
    if (tets && tetCount > 0)
        for (int t = 0; t < tetCount; ++t)
        {
            const int* r = tets + 4 * t;
            bool bad = false;
            for (int a = 0; a < 4; ++a)
                bad = bad || r[a] < 0 || r[a] >= nodeCount;
            if (bad)
                continue;
            const int a = remap[r[0]], b = remap[r[1]], c = remap[r[2]], d = remap[r[3]];
            if (a == b || a == c || a == d || b == c || b == d || c == d)
                continue;
            const S* pa = kept.data() + 3 * a;
            const S* pb = kept.data() + 3 * b;
            const S* pc = kept.data() + 3 * c;
            const S* pd = kept.data() + 3 * d;
            const S ab[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
            const S ac[3] = {pc[0] - pa[0], pc[1] - pa[1], pc[2] - pa[2]};
            const S ad[3] = {pd[0] - pa[0], pd[1] - pa[1], pd[2] - pa[2]};
            const S cx = ab[1] * ac[2] - ab[2] * ac[1], cy = ab[2] * ac[0] - ab[0] * ac[2],
                    cz = ab[0] * ac[1] - ab[1] * ac[0];
            const S six = cx * ad[0] + cy * ad[1] + cz * ad[2];
            if (six * six <= static_cast<S>(1.0e-24f))
                continue;
            m->tets.push_back(a);
            m->tets.push_back(b);
            m->tets.push_back(c);
            m->tets.push_back(d);
        }
    return m;
}

struct Component
{
    bool physics = false;
    bool active = true;
    int globalBegin = -1;
    int globalCount = 0;
    std::vector<double> rest;   // local rest positions
    std::vector<int> tets;      // local ids
    std::vector<uint8_t> tetActive;
    std::vector<double> young;  // per tet
    Material material;
    int femModel = 1;  // TetMeshPhysicsComponentDesc::femModel: 0 Linear, 1 Corotational, 2 NeoHookean
    int tetBase = -1;  // index of the component's first tet in the device system
    // the component's own copy of its current node positions
    // (TetMeshComponent nodePositions: the rest positions at creation,
    // PHYSIK_SetTetMeshLocalCurrentPosition writes it); a physics component
    // copies the world's positions into it in every substep and after a step
    // with dt > 0 (TetMeshPhysicsComponent.cpp:234-252, 299, 356-362).
    // currentFromWorld: the copy equals the world's present positions and is
    // read from there until the world state changes on the host.
    std::vector<float> current;
    bool currentFromWorld = false;
    // a collision sphere instead of a tet mesh (CollisionSphereComponent.h,
    // CollisionComponent.h:10-28): centre, radius clamped at 0, connection
    // settings (defaults 1000 and 10), and a kinematic target taken at the
    // next step
    bool sphere = false;
    float center[3] = {0.f, 0.f, 0.f};
    float radius = 0.f;
    float stiffness = 1000.f, damping = 10.f;
    bool hasTarget = false;
    float target[3] = {0.f, 0.f, 0.f};
};

struct World
{
    std::vector<double> x, v;
    std::vector<uint8_t> fixed, nodeActive;
    std::vector<std::unique_ptr<Component>> components;
    std::vector<uint32_t> generations;
    std::vector<uint32_t> freeSlots;

    float gravity[3] = {0.f, 0.f, 0.f};
    int substeps = 1;
    bool implicit = false;  // SolverMode::Explicit by default (World.h:133)
    float tolerance = 1.0e-4f;
    int maxIterations = 128;
    int precision = PK_FP64;
    int lastIterations = 0;
    float lastResidual = 0.f;
    bool lastConverged = false;
    int lastStatus = PK_GROUP_NO_DYNAMIC;

    // point connections added for the next frame (World.cpp:126-131), and the
    // node pairs they have needed in the BSR pattern beyond the tets' (the
    // system is rebuilt with them, SolverData.cpp:389-405)
    struct Conn
    {
        int n[4];
        double w[4], target[3], k, c;
    };
    std::vector<Conn> conns;
    std::set<std::pair<int, int>> extraPairs;
    bool sysSpheres = false;  // the device system holds collision spheres
    int builds = 0;           // device systems built so far

    // tets switched since the last step that the device system can take in
    // place (system tet id, new flag), and per node the number of active tets
    // of active physics components that reference it (built on demand)
    std::vector<std::pair<int, uint8_t>> pendingTets;
    std::vector<int> refs;
    bool refsValid = false;

    pk_system* sys = nullptr;
    bool topoDirty = true;     // rebuild the device system before the next step
    bool hostNewer = false;    // host state changed since the last upload
    bool deviceNewer = false;  // device state changed since the last download
    // frame export: what each step exports for the float getters (default
    // both), what the last step did export, and the float arrays once waited on
    int exportMask = PK_EXPORT_POSITIONS | PK_EXPORT_VELOCITIES;
    int frameMask = 0;
    bool frameWaited = false;
    const float* frameX = nullptr;
    const float* frameV = nullptr;

    ~World()
    {
        if (sys)
            pk_system_destroy(sys);
    }
    int nodeCount() const { return static_cast<int>(fixed.size()); }

    void pull()
    {
        if (!deviceNewer || !sys)
            return;
        if (pk_system_get_state(sys, x.data(), v.data()) == PK_OK)
            deviceNewer = false;
    }
    // The float node state of the last step's frame export (`which` is
    // PK_EXPORT_POSITIONS or PK_EXPORT_VELOCITIES) while the device holds the
    // newest state; nullptr sends the getter to the host mirror (pull()).
    // Each value is static_cast<float> of the FP64 state, as from the mirror.
    const float* frame(int which)
    {
        if (!deviceNewer || !sys || !(frameMask & which))
            return nullptr;
        if (!frameWaited)
        {
            const bool px = (frameMask & PK_EXPORT_POSITIONS) != 0, pv = (frameMask & PK_EXPORT_VELOCITIES) != 0;
            if (pk_system_export_view(sys, px ? &frameX : nullptr, pv ? &frameV : nullptr) != PK_OK)
            {
                frameMask = 0;
                return nullptr;
            }
            frameWaited = true;
        }
        return which == PK_EXPORT_POSITIONS ? frameX : frameV;
    }
    void touch_host()
    {
        pull();
        materialize_current();
        hostNewer = true;
    }
    // float world position of node g, as PHYSIK_GetNodePosition returns it
    void node_position(int g, float* out)
    {
        if (const float* f = frame(PK_EXPORT_POSITIONS))
        {
            std::copy(f + 3 * g, f + 3 * g + 3, out);
            return;
        }
        pull();
        for (int k = 0; k < 3; ++k)
            out[k] = static_cast<float>(x[3 * g + k]);
    }
    // The copies that still read the world take its positions now, before
    // the host changes the world state.
    void materialize_current(Component& c)
    {
        if (!c.currentFromWorld)
            return;
        c.currentFromWorld = false;
        for (int i = 0; i < c.globalCount && 3 * i + 2 < static_cast<int>(c.current.size()); ++i)
            if (c.globalBegin + i < nodeCount())
                node_position(c.globalBegin + i, c.current.data() + 3 * i);
    }
    void materialize_current()
    {
        for (auto& cp : components)
            if (cp)
                materialize_current(*cp);
    }
    int add_node(double px, double py, double pz)
    {
        touch_host();
        x.push_back(px);
        x.push_back(py);
        x.push_back(pz);
        v.push_back(0.0);
        v.push_back(0.0);
        v.push_back(0.0);
        fixed.push_back(0);
        nodeActive.push_back(1);
        topoDirty = true;
        return nodeCount() - 1;
    }
    Component* get(PhysiK::ComponentHandle h)
    {
        if (!h.IsValid() || h.index >= components.size())
            return nullptr;
        if (generations[h.index] != h.generation || !components[h.index])
            return nullptr;
        return components[h.index].get();
    }
    PhysiK::ComponentHandle add(std::unique_ptr<Component> c)  // World.cpp:448-481
    {
        refsValid = false;
        if (!freeSlots.empty())
        {
            const uint32_t slot = freeSlots.back();
            freeSlots.pop_back();
            components[slot] = std::move(c);
            return PhysiK::ComponentHandle{slot, generations[slot]};
        }
        components.push_back(std::move(c));
        generations.push_back(1u);
        return PhysiK::ComponentHandle{static_cast<uint32_t>(components.size() - 1), 1u};
    }
    void destroy(PhysiK::ComponentHandle h)  // World.cpp:105-124
    {
        if (!get(h))
            return;
        const bool sphere = components[h.index]->sphere;
        refsValid = false;
        components[h.index].reset();
        ++generations[h.index];
        if (generations[h.index] == 0u)
            generations[h.index] = 1u;
        freeSlots.push_back(h.index);
        if (!sphere)
            topoDirty = true;  // spheres reach the system at every step
    }

    // Rebuilds the device system from every active physics component.
    int rebuild()
    {
        pull();
        const int N = nodeCount();
        std::vector<int> tets;
        std::vector<double> young, poisson, damping, density;
        std::vector<uint8_t> tact, models;
        bool mixed = false;
        pendingTets.clear();
        for (const auto& cp : components)
        {
            if (!cp || !cp->physics || !cp->active)
                continue;
            Component& c = *cp;
            c.tetBase = static_cast<int>(young.size());
            mixed = mixed || c.femModel != 1;
            for (size_t t = 0; t < c.tetActive.size(); ++t)
            {
                models.push_back(static_cast<uint8_t>(c.femModel));
                for (int a = 0; a < 4; ++a)
                    tets.push_back(c.globalBegin + c.tets[4 * t + a]);
                young.push_back(c.young[t]);
                poisson.push_back(c.material.poissonRatio);
                damping.push_back(c.material.damping);
                density.push_back(c.material.density);
                tact.push_back(c.tetActive[t]);
            }
        }
        pk_mesh_desc d{};
        d.node_count = N;
        d.rest = x.data();  // rest data is per tet; the system's x is replaced below
        d.tet_count = static_cast<int>(young.size());
        d.tets = tets.data();
        d.young = young.data();
        d.poisson = poisson.data();
        d.damping = damping.data();
        d.density = density.data();
        d.tet_active = tact.data();
        d.node_fixed = fixed.data();
        d.node_active = nodeActive.data();
        std::vector<int> extra;
        for (const auto& pr : extraPairs)
            if (pr.first < N && pr.second < N)
            {
                extra.push_back(pr.first);
                extra.push_back(pr.second);
            }
        d.extra_pair_count = static_cast<int>(extra.size() / 2);
        d.extra_pairs = extra.empty() ? nullptr : extra.data();
        // Rest geometry comes from each component's own rest positions.
        std::vector<double> rest(x);
        for (const auto& cp : components)
            if (cp && cp->physics && cp->active)
                std::copy(cp->rest.begin(), cp->rest.end(), rest.begin() + 3 * cp->globalBegin);
        d.rest = rest.data();
        if (sys)
        {
            pk_system_destroy(sys);
            sys = nullptr;
        }
        int rc = pk_system_create(&d, precision, &sys);
        if (rc != PK_OK)
            return rc;
        ++builds;
        if (mixed && (rc = pk_system_set_fem_models(sys, models.data())) != PK_OK)
            return rc;
        rc = pk_system_set_state(sys, x.data(), v.data());
        if (rc != PK_OK)
            return rc;
        topoDirty = false;
        hostNewer = false;
        deviceNewer = false;
        sysSpheres = false;
        return PK_OK;
    }

    // Hands the collision spheres to the system, in component order. A sphere
    // of radius 0 touches nothing (CollisionSphereComponent.cpp:120-123), nor
    // does one with a non-finite centre or radius: every contact it could make
    // has a non-finite target, which PointConnection::UpdateSystem skips.
    int apply_spheres()
    {
        std::vector<double> c, r, k, d;
        for (const auto& cp : components)
        {
            if (!cp || !cp->sphere || !cp->active || !(cp->radius > 0.f) || !std::isfinite(cp->radius))
                continue;
            if (!std::isfinite(cp->center[0]) || !std::isfinite(cp->center[1]) || !std::isfinite(cp->center[2]))
                continue;
            c.insert(c.end(), cp->center, cp->center + 3);
            r.push_back(cp->radius);
            k.push_back(cp->stiffness);
            d.push_back(cp->damping);
        }
        if (r.empty() && !sysSpheres)
            return PK_OK;
        const int rc = pk_system_set_spheres(sys, static_cast<int>(r.size()), c.data(), r.data(), k.data(), d.data());
        if (rc == PK_OK)
            sysSpheres = !r.empty();
        return rc;
    }

    // Hands the frame's connections to the system; when they couple nodes the
    // pattern lacks, rebuilds it with those pairs first.
    int apply_connections()
    {
        auto upload = [&]() {
            std::vector<int> n;
            std::vector<double> w, tg, k, c;
            for (const Conn& cn : conns)
            {
                n.insert(n.end(), cn.n, cn.n + 4);
                w.insert(w.end(), cn.w, cn.w + 4);
                tg.insert(tg.end(), cn.target, cn.target + 3);
                k.push_back(cn.k);
                c.push_back(cn.c);
            }
            return pk_system_set_connections(sys, static_cast<int>(conns.size()), n.data(), w.data(), tg.data(),
                                             k.data(), c.data());
        };
        int rc = upload();
        if (rc != PK_ERR_ARG)
            return rc;
        for (const Conn& cn : conns)
            for (int a = 0; a < 4; ++a)
                for (int b = a + 1; b < 4; ++b)
                    if (cn.w[a] != 0.0 && cn.w[b] != 0.0 && cn.n[a] != cn.n[b])
                        extraPairs.insert({std::min(cn.n[a], cn.n[b]), std::max(cn.n[a], cn.n[b])});
        if ((rc = rebuild()) != PK_OK)
            return rc;
        return upload();
    }

    int step(float frameDt)  // World.cpp:32-73
    {
        if (frameDt < 0.0f)
            return PK_OK;
        // World::PreUpdateComponents (World.cpp:38, 527-537): a collision
        // component takes its kinematic target (CollisionComponent.cpp:5-15)
        for (auto& cp : components)
            if (cp && cp->sphere && cp->active && cp->hasTarget)
            {
                std::copy(cp->target, cp->target + 3, cp->center);
                cp->hasTarget = false;
            }
        if (frameDt == 0.0f)
        {
            conns.clear();  // World.cpp:41-46
            return PK_OK;
        }
        // every active physics component copies the world's positions in its
        // first UpdateSystem and again in PostUpdate: from here on its copy
        // is the world's state (also when the step throws after the copy)
        for (auto& cp : components)
            if (cp && cp->physics && cp->active)
                cp->currentFromWorld = true;
        if (!topoDirty && sys && !pendingTets.empty())
        {
            // tets switched since the last step, in place when the system can
            // take them (TetMeshComponent.cpp:153-169)
            std::vector<int> ids;
            std::vector<unsigned char> flags;
            for (const auto& p : pendingTets)
            {
                ids.push_back(p.first);
                flags.push_back(p.second);
            }
            const int rc = pk_system_set_tets_active(sys, static_cast<int>(ids.size()), ids.data(), flags.data());
            if (rc == PK_ERR_REBUILD)
                topoDirty = true;
            else if (rc != PK_OK)
                return rc;
            pendingTets.clear();
        }
        if (topoDirty || !sys)
        {
            const int rc = rebuild();
            if (rc != PK_OK)
                return rc;
        }
        if (hostNewer)
        {
            const int rc = pk_system_set_state(sys, x.data(), v.data());
            if (rc != PK_OK)
                return rc;
            hostNewer = false;
        }
        // the connections first: a pattern they need rebuilds the system,
        // which the spheres and settings below must then reach
        if (!conns.empty())
        {
            const int rc = apply_connections();
            if (rc != PK_OK)
                return rc;
        }
        if (const int rc = apply_spheres(); rc != PK_OK)
            return rc;
        const int steps = std::max(1, substeps);
        pk_system_set_gravity(sys, gravity[0], gravity[1], gravity[2]);
        pk_system_set_mode(sys, implicit ? 1 : 0);
        int rc = pk_system_set_solver(sys, tolerance, maxIterations, 1);
        if (rc != PK_OK)
            return rc;
        frameMask = 0;
        frameWaited = false;
        rc = pk_system_step(sys, static_cast<double>(frameDt), steps, 1);
        if (!conns.empty())
            pk_system_set_connections(sys, 0, nullptr, nullptr, nullptr, nullptr, nullptr);
        if (rc != PK_OK)
            return rc;  // the reference leaves the frame's connections when it throws
        conns.clear();  // World::ClearFrameConnections (World.cpp:70)
        deviceNewer = true;
        // the readback for the frame's readers starts now and overlaps the
        // host's work until the first getter waits for it
        if (exportMask && pk_system_export_begin(sys, exportMask) == PK_OK)
            frameMask = exportMask;
        if (implicit)
        {
            int it = 0, conv = 0, status = 0;
            double res = 0.0;
            pk_system_group_results(sys, &it, &res, &conv, &status);
            lastIterations = it;
            lastResidual = static_cast<float>(res);
            lastConverged = conv != 0;
            lastStatus = status;
        }
        return PK_OK;
    }
};

World* W(PhysiK::WorldHandle h) { return static_cast<World*>(h); }

GeneratedMesh* M(PhysiK::GeneratedTetMeshHandle h)
{
    auto* m = static_cast<GeneratedMesh*>(h);
    return (m && g_meshes.count(m)) ? m : nullptr;
}

bool valid_node(World* w, int i) { return w && i >= 0 && i < w->nodeCount(); }

Component* tet_mesh(World* w, PhysiK::ComponentHandle h)
{
    Component* c = w ? w->get(h) : nullptr;
    return (c && !c->sphere) ? c : nullptr;
}

Component* sphere_of(World* w, PhysiK::ComponentHandle h)  // AsCollisionSphere (PhysiKAPI.cpp:101-112)
{
    Component* c = w ? w->get(h) : nullptr;
    return (c && c->sphere) ? c : nullptr;
}

// SanitizeConnectionValue (CollisionSphereComponent.cpp:38-46)
float sanitize_connection_value(float value) { return std::isfinite(value) ? std::max(0.0f, value) : 0.0f; }

// CollisionSphereComponent::QueryOverlaps (CollisionSphereComponent.cpp:195-285):
// every active tet of an active tet-mesh physics component, in component
// order, with at least one node within the radius, in float as the reference
// computes it from its float positions.
std::vector<PhysikCollisionSphereOverlap> query_overlaps(World* w, PhysiK::ComponentHandle h)
{
    std::vector<PhysikCollisionSphereOverlap> out;
    const Component* s = sphere_of(w, h);
    if (!s || !s->active || s->radius <= 0.0f)
        return out;
    w->pull();
    const float r2 = s->radius * s->radius;
    for (size_t ci = 0; ci < w->components.size(); ++ci)
    {
        const auto& cp = w->components[ci];
        if (!cp || !cp->physics || !cp->active)
            continue;
        for (size_t t = 0; t < cp->tetActive.size(); ++t)
        {
            if (!cp->tetActive[t])
                continue;
            int nodes[4];
            bool ok = true;
            for (int a = 0; a < 4; ++a)
            {
                nodes[a] = cp->globalBegin + cp->tets[4 * t + a];
                ok = ok && valid_node(w, nodes[a]);
            }
            if (!ok)
                continue;
            int mask = 0, count = 0;
            float minDistance = 0.0f;
            bool hasDistance = false;
            for (int a = 0; a < 4; ++a)
            {
                const float dx = static_cast<float>(w->x[3 * nodes[a]]) - s->center[0];
                const float dy = static_cast<float>(w->x[3 * nodes[a] + 1]) - s->center[1];
                const float dz = static_cast<float>(w->x[3 * nodes[a] + 2]) - s->center[2];
                const float xx = dx * dx, yy = dy * dy, zz = dz * dz;
                const float dsq = (xx + yy) + zz;
                const float distance = std::sqrt(std::max(0.0f, dsq));
                if (!hasDistance || distance < minDistance)
                {
                    minDistance = distance;
                    hasDistance = true;
                }
                if (dsq <= r2)
                {
                    mask |= 1 << a;
                    ++count;
                }
            }
            if (count <= 0)
                continue;
            PhysikCollisionSphereOverlap o{};
            o.geometryType = PHYSIK_OverlapGeometry_Tetrahedron;
            o.component = PhysiK::ComponentHandle{static_cast<std::uint32_t>(ci), w->generations[ci]};
            o.primitiveIndex = static_cast<int>(t);
            o.node0 = nodes[0];
            o.node1 = nodes[1];
            o.node2 = nodes[2];
            o.node3 = nodes[3];
            o.overlappedNodeMask = mask;
            o.overlappedNodeCount = count;
            o.sphereCenterX = s->center[0];
            o.sphereCenterY = s->center[1];
            o.sphereCenterZ = s->center[2];
            o.sphereRadius = s->radius;
            o.minDistance = minDistance;
            out.push_back(o);
        }
    }
    return out;
}

Component* physics(World* w, PhysiK::ComponentHandle h)
{
    Component* c = tet_mesh(w, h);
    return (c && c->physics) ? c : nullptr;
}

int global_of(const Component& c, int local)  // TetMeshPhysicsComponent.cpp:264-274
{
    if (!c.physics || c.globalBegin < 0 || local < 0 || local >= c.globalCount)
        return -1;
    return c.globalBegin + local;
}

// PhysiKAPI.cpp:159-229: a node stays active while an active tet of an active
// tet-mesh component references it; a deactivated node loses its velocity.
// Per node, how many corners of active tets of active tet-mesh physics
// components reference it: ActiveTetReferencesGlobalNode (PhysiKAPI.cpp:158-203)
// is "count > 0". Counted once, then kept up to date by PHYSIK_SetTetActive.
void ensure_refs(World* w)
{
    if (w->refsValid && w->refs.size() == static_cast<size_t>(w->nodeCount()))
        return;
    w->refs.assign(static_cast<size_t>(w->nodeCount()), 0);
    for (const auto& cp : w->components)
    {
        if (!cp || !cp->active || !cp->physics)
            continue;
        for (size_t t = 0; t < cp->tetActive.size(); ++t)
            if (cp->tetActive[t])
                for (int b = 0; b < 4; ++b)
                {
                    const int g = global_of(*cp, cp->tets[4 * t + b]);
                    if (valid_node(w, g))
                        ++w->refs[g];
                }
    }
    w->refsValid = true;
}

std::unique_ptr<Component> component_from(const GeneratedMesh& m)  // TetMeshComponent.cpp:31-67
{
    auto c = std::make_unique<Component>();
    c->rest = m.positions;
    const int n = static_cast<int>(m.positions.size() / 3);
    for (size_t t = 0; t + 3 < m.tets.size(); t += 4)
    {
        bool ok = true;
        for (int a = 0; a < 4; ++a)
            ok = ok && m.tets[t + a] >= 0 && m.tets[t + a] < n;
        if (!ok)
            continue;
        c->tets.insert(c->tets.end(), m.tets.begin() + t, m.tets.begin() + t + 4);
    }
    c->tetActive.assign(c->tets.size() / 4, 1);
    c->current.assign(m.positions.begin(), m.positions.end());
    return c;
}

void apply_material(Component& c, const Material& m)
{
    c.material = m;
    c.young.assign(c.tetActive.size(), m.youngModulus);
}

Material to_material(const PhysikMaterialDesc& d)  // PhysiKAPI.cpp:50-58
{
    Material m;
    m.density = d.density;
    m.youngModulus = d.youngModulus;
    m.poissonRatio = d.poissonRatio;
    m.damping = d.damping;
    return m;
}

}  // namespace physik_amd

using namespace physik_amd;

extern "C"
{

PhysiK::WorldHandle PHYSIK_CreateWorld() { return new World(); }

void PHYSIK_DestroyWorld(PhysiK::WorldHandle world) { delete W(world); }

void PHYSIK_Step(PhysiK::WorldHandle world, float dt)
{
    World* w = W(world);
    if (!w)
        return;
    const int rc = w->step(dt);
    if (rc == PK_ERR_MASS)
        throw std::runtime_error(pk_last_error());
    if (rc != PK_OK)
        throw std::runtime_error(std::string("PHYSIK_Step: ") + pk_last_error());
}

int PHYSIK_AMD_TryStep(PhysiK::WorldHandle world, float dt)
{
    World* w = W(world);
    if (!w)
        return 1;
    const int rc = w->step(dt);
    if (rc == PK_OK)
        return 0;
    g_apiError = pk_last_error();
    return rc == PK_ERR_MASS ? 3 : 2;
}

void PHYSIK_SetSubstepCount(PhysiK::WorldHandle world, int substepCount)
{
    if (World* w = W(world))
        w->substeps = std::max(1, substepCount);
}

void PHYSIK_SetSolverMode(PhysiK::WorldHandle world, int mode)
{
    if (World* w = W(world))
        w->implicit = mode == 1;
}

void PHYSIK_SetConjugateGradientTolerance(PhysiK::WorldHandle world, float tolerance)
{
    World* w = W(world);
    if (!w || !std::isfinite(tolerance))
        return;
    w->tolerance = std::clamp(tolerance, 1.0e-8f, 1.0e-1f);  // World.cpp:177-188
}

float PHYSIK_GetConjugateGradientTolerance(PhysiK::WorldHandle world)
{
    World* w = W(world);
    return w ? w->tolerance : 0.0f;
}

void PHYSIK_SetConjugateGradientMaxIterations(PhysiK::WorldHandle world, int maxIterations)
{
    if (World* w = W(world))
        w->maxIterations = std::clamp(maxIterations, 1, 1024);  // World.cpp:195-201
}

int PHYSIK_GetConjugateGradientMaxIterations(PhysiK::WorldHandle world)
{
    World* w = W(world);
    return w ? w->maxIterations : 0;
}

int PHYSIK_GetLastConjugateGradientIterations(PhysiK::WorldHandle world)
{
    World* w = W(world);
    return w ? w->lastIterations : 0;
}

float PHYSIK_GetLastConjugateGradientResidualNorm(PhysiK::WorldHandle world)
{
    World* w = W(world);
    return w ? w->lastResidual : 0.0f;
}

int PHYSIK_DidLastConjugateGradientSolveConverge(PhysiK::WorldHandle world)
{
    World* w = W(world);
    return (w && w->lastConverged) ? 1 : 0;
}

void PHYSIK_SetGravity(PhysiK::WorldHandle world, float x, float y, float z)
{
    if (World* w = W(world))
    {
        w->gravity[0] = x;
        w->gravity[1] = y;
        w->gravity[2] = z;
    }
}

int PHYSIK_AddNode(PhysiK::WorldHandle world, float x, float y, float z)
{
    World* w = W(world);
    return w ? w->add_node(x, y, z) : -1;
}

void PHYSIK_SetNodeFixed(PhysiK::WorldHandle world, int nodeIndex, int fixed)  // World.cpp:403-417
{
    World* w = W(world);
    if (!valid_node(w, nodeIndex))
        return;
    w->touch_host();
    const uint8_t f = fixed != 0 ? 1 : 0;
    if (f)
        for (int i = 0; i < 3; ++i)
            w->v[3 * nodeIndex + i] = 0.0;
    if (w->fixed[nodeIndex] != f)
        w->topoDirty = true;
    w->fixed[nodeIndex] = f;
}

int PHYSIK_IsNodeFixed(PhysiK::WorldHandle world, int nodeIndex)
{
    World* w = W(world);
    return (valid_node(w, nodeIndex) && w->fixed[nodeIndex]) ? 1 : 0;
}

PhysiK::GeneratedTetMeshHandle PHYSIK_GenerateTetMesh(const PhysiK::Vec3* positions, int nodeCount,
                                                      const int* tetLocalNodeIndices, int tetCount)
{
    GeneratedMesh* m = generate<float>(reinterpret_cast<const float*>(positions), nodeCount, tetLocalNodeIndices,
                                       tetCount);
    g_meshes.insert(m);
    return m;
}

PhysiK::GeneratedTetMeshHandle PHYSIK_AMD_GenerateTetMeshF64(const double* positions, int nodeCount,
                                                             const int* tetLocalNodeIndices, int tetCount)
{
    GeneratedMesh* m = generate<double>(positions, nodeCount, tetLocalNodeIndices, tetCount);
    g_meshes.insert(m);
    return m;
}

int PHYSIK_IsGeneratedTetMeshHandleValid(PhysiK::GeneratedTetMeshHandle h) { return M(h) ? 1 : 0; }

void PHYSIK_DestroyGeneratedTetMesh(PhysiK::GeneratedTetMeshHandle h)
{
    GeneratedMesh* m = M(h);
    if (!m)
        return;
    g_meshes.erase(m);
    delete m;
}

int PHYSIK_GetGeneratedTetMeshVertexCount(PhysiK::GeneratedTetMeshHandle h)
{
    GeneratedMesh* m = M(h);
    return m ? static_cast<int>(m->positions.size() / 3) : 0;
}

int PHYSIK_GetGeneratedTetMeshTetCount(PhysiK::GeneratedTetMeshHandle h)
{
    GeneratedMesh* m = M(h);
    return m ? static_cast<int>(m->tets.size() / 4) : 0;
}

int PHYSIK_GetGeneratedTetMeshTetIndexCount(PhysiK::GeneratedTetMeshHandle h)
{
    GeneratedMesh* m = M(h);
    return m ? static_cast<int>(m->tets.size()) : 0;
}

int PHYSIK_GetGeneratedTetMeshVertex(PhysiK::GeneratedTetMeshHandle h, int i, float* outX, float* outY, float* outZ)
{
    GeneratedMesh* m = M(h);
    if (!m || !outX || !outY || !outZ || i < 0 || i >= static_cast<int>(m->positions.size() / 3))
        return 0;
    *outX = static_cast<float>(m->positions[3 * i]);
    *outY = static_cast<float>(m->positions[3 * i + 1]);
    *outZ = static_cast<float>(m->positions[3 * i + 2]);
    return 1;
}

int PHYSIK_GetGeneratedTetMeshTetNodeIndex(PhysiK::GeneratedTetMeshHandle h, int i, int* outNodeIndex)
{
    GeneratedMesh* m = M(h);
    if (!m || !outNodeIndex || i < 0 || i >= static_cast<int>(m->tets.size()))
        return 0;
    *outNodeIndex = m->tets[i];
    return 1;
}

PhysiK::ComponentHandle PHYSIK_CreateTetMeshComponent(PhysiK::WorldHandle world, PhysiK::GeneratedTetMeshHandle h)
{
    World* w = W(world);
    GeneratedMesh* m = M(h);
    if (!w || !m || m->positions.empty() || m->tets.empty())
        return PhysiK::ComponentHandle{};
    return w->add(component_from(*m));
}

// TetMeshPhysicsComponent::CreateFromGeneratedTetMesh (TetMeshPhysicsComponent.cpp:163-187)
PhysiK::ComponentHandle PHYSIK_CreateTetMeshPhysicsComponent(PhysiK::WorldHandle world,
                                                             PhysiK::GeneratedTetMeshHandle h,
                                                             const PhysikMaterialDesc* material)
{
    World* w = W(world);
    GeneratedMesh* m = M(h);
    if (!w || !m || !material || m->positions.empty() || m->tets.empty())
        return PhysiK::ComponentHandle{};
    auto c = component_from(*m);
    c->physics = true;
    apply_material(*c, to_material(*material));
    c->globalBegin = w->nodeCount();
    c->globalCount = static_cast<int>(c->rest.size() / 3);
    for (int i = 0; i < c->globalCount; ++i)
        w->add_node(c->rest[3 * i], c->rest[3 * i + 1], c->rest[3 * i + 2]);
    w->topoDirty = true;
    return w->add(std::move(c));
}

int PHYSIK_SetTetMeshLocalCurrentPosition(PhysiK::WorldHandle world, PhysiK::ComponentHandle h, int local, float x,
                                          float y, float z)  // PhysiKAPI.cpp:992-1016
{
    World* w = W(world);
    Component* c = tet_mesh(w, h);
    if (!c || local < 0 || 3 * local + 2 >= static_cast<int>(c->current.size()))
        return 0;
    w->materialize_current(*c);
    c->current[3 * local] = x;
    c->current[3 * local + 1] = y;
    c->current[3 * local + 2] = z;
    return 1;
}

int PHYSIK_GetTetMeshLocalCurrentPosition(PhysiK::WorldHandle world, PhysiK::ComponentHandle h, int local,
                                          float* outX, float* outY, float* outZ)  // PhysiKAPI.cpp:1018-1053
{
    World* w = W(world);
    if (!w || !outX || !outY || !outZ)
        return 0;
    Component* c = tet_mesh(w, h);
    if (!c || local < 0 || 3 * local + 2 >= static_cast<int>(c->current.size()))
        return 0;
    float p[3] = {c->current[3 * local], c->current[3 * local + 1], c->current[3 * local + 2]};
    if (c->currentFromWorld && local < c->globalCount && c->globalBegin + local < w->nodeCount())
        w->node_position(c->globalBegin + local, p);
    *outX = p[0];
    *outY = p[1];
    *outZ = p[2];
    return 1;
}

int PHYSIK_GetTetMeshGlobalNodeBeginIndex(PhysiK::WorldHandle world, PhysiK::ComponentHandle h)
{
    Component* c = physics(W(world), h);
    return c ? c->globalBegin : -1;
}

int PHYSIK_GetTetMeshGlobalNodeCount(PhysiK::WorldHandle world, PhysiK::ComponentHandle h)
{
    Component* c = physics(W(world), h);
    return c ? c->globalCount : -1;
}

int PHYSIK_GetTetMeshGlobalNodeIndex(PhysiK::WorldHandle world, PhysiK::ComponentHandle h, int local)
{
    Component* c = physics(W(world), h);
    return c ? global_of(*c, local) : -1;
}

void PHYSIK_SetTetMeshMaterial(PhysiK::WorldHandle world, PhysiK::ComponentHandle h,
                               const PhysikMaterialDesc* material)
{
    World* w = W(world);
    if (!w || !material)
        return;
    Component* c = physics(w, h);
    if (!c)
        return;
    apply_material(*c, to_material(*material));
    w->topoDirty = true;
}

int PHYSIK_AMD_SetTetMeshYoungModuli(PhysiK::WorldHandle world, PhysiK::ComponentHandle h, const double* young,
                                     int count)
{
    World* w = W(world);
    Component* c = physics(w, h);
    if (!c || !young || count != static_cast<int>(c->young.size()))
        return 0;
    c->young.assign(young, young + count);
    w->topoDirty = true;
    return 1;
}

int PHYSIK_AMD_SetTetMeshFemModel(PhysiK::WorldHandle world, PhysiK::ComponentHandle h, int femModel)
{
    World* w = W(world);
    Component* c = physics(w, h);
    if (!c || femModel < 0 || femModel > 2)
        return 0;
    if (c->femModel != femModel)
        w->topoDirty = true;
    c->femModel = femModel;
    return 1;
}

void PHYSIK_DestroyComponent(PhysiK::WorldHandle world, PhysiK::ComponentHandle h)
{
    if (World* w = W(world))
        w->destroy(h);
}

int PHYSIK_IsComponentHandleValid(PhysiK::WorldHandle world, PhysiK::ComponentHandle h)
{
    World* w = W(world);
    return (w && w->get(h)) ? 1 : 0;
}

int PHYSIK_GetTetMeshTetCount(PhysiK::WorldHandle world, PhysiK::ComponentHandle h)
{
    Component* c = tet_mesh(W(world), h);
    return c ? static_cast<int>(c->tetActive.size()) : 0;
}

int PHYSIK_GetTetMeshNodeCount(PhysiK::WorldHandle world, PhysiK::ComponentHandle h)
{
    Component* c = tet_mesh(W(world), h);
    return c ? static_cast<int>(c->rest.size() / 3) : 0;
}

int PHYSIK_IsTetActive(PhysiK::WorldHandle world, PhysiK::ComponentHandle h, int tet)
{
    Component* c = tet_mesh(W(world), h);
    if (!c || tet < 0 || tet >= static_cast<int>(c->tetActive.size()))
        return 0;
    return c->tetActive[tet] ? 1 : 0;
}

void PHYSIK_SetTetActive(PhysiK::WorldHandle world, PhysiK::ComponentHandle h, int tet, int active)
{
    World* w = W(world);
    Component* c = tet_mesh(w, h);
    if (!c || tet < 0 || tet >= static_cast<int>(c->tetActive.size()))
        return;
    const uint8_t a = active != 0 ? 1 : 0;
    if (c->tetActive[tet] == a)
        return;
    ensure_refs(w);
    c->tetActive[tet] = a;
    int g[4];
    for (int b = 0; b < 4; ++b)
    {
        g[b] = global_of(*c, c->tets[4 * tet + b]);
        if (c->physics && c->active && valid_node(w, g[b]))
            w->refs[g[b]] += a ? 1 : -1;
    }
    // ReconcileTetNodeActiveStates (PhysiKAPI.cpp:206-229): a corner node
    // stays active while an active tet references it; one that is not loses
    // its velocity
    bool changed = false, unreferenced = false;
    for (int b = 0; b < 4; ++b)
        if (valid_node(w, g[b]))
        {
            const uint8_t ref = w->refs[g[b]] > 0 ? 1 : 0;
            changed = changed || w->nodeActive[g[b]] != ref;
            unreferenced = unreferenced || !ref;
        }
    if (changed || unreferenced)
    {
        w->touch_host();
        for (int b = 0; b < 4; ++b)
            if (valid_node(w, g[b]))
            {
                w->nodeActive[g[b]] = w->refs[g[b]] > 0 ? 1 : 0;
                if (!w->nodeActive[g[b]])
                    for (int i = 0; i < 3; ++i)
                        w->v[3 * g[b] + i] = 0.0;
            }
    }
    if (!c->physics)
        return;
    // the device system takes the switch in place at the next step unless the
    // set of active nodes changed (it then falls back to building it again)
    if (changed || !w->sys || w->topoDirty || c->tetBase < 0)
        w->topoDirty = true;
    else
        w->pendingTets.push_back({c->tetBase + tet, a});
}

int PHYSIK_AMD_GetSystemBuildCount(PhysiK::WorldHandle world)
{
    World* w = W(world);
    return w ? w->builds : 0;
}

void PHYSIK_DeactivateTet(PhysiK::WorldHandle world, PhysiK::ComponentHandle h, int tet)
{
    PHYSIK_SetTetActive(world, h, tet, 0);
}

int PHYSIK_GetActiveTetCount(PhysiK::WorldHandle world, PhysiK::ComponentHandle h)
{
    Component* c = tet_mesh(W(world), h);
    if (!c)
        return 0;
    return static_cast<int>(std::count(c->tetActive.begin(), c->tetActive.end(), 1));
}

// PhysiKAPI.cpp:685-700, CollisionSphereComponent::Create (CollisionSphereComponent.cpp:69-77)
PhysiK::ComponentHandle PHYSIK_CreateCollisionSphereComponent(PhysiK::WorldHandle world, float x, float y, float z,
                                                              float radius)
{
    World* w = W(world);
    if (!w)
        return PhysiK::ComponentHandle{};
    auto c = std::make_unique<Component>();
    c->sphere = true;
    c->center[0] = x;
    c->center[1] = y;
    c->center[2] = z;
    c->radius = std::max(0.0f, radius);
    return w->add(std::move(c));
}

// PhysiKAPI.cpp:1180-1194 -> CollisionSphereComponent::SetConnectionSettings (:86-105)
void PHYSIK_SetCollisionSphereConnectionSettings(PhysiK::WorldHandle world, PhysiK::ComponentHandle h,
                                                 float stiffness, float damping)
{
    Component* c = sphere_of(W(world), h);
    if (!c)
        return;
    c->stiffness = sanitize_connection_value(stiffness);
    c->damping = sanitize_connection_value(damping);
}

// PhysiKAPI.cpp:1196-1216
void PHYSIK_GetCollisionSphereConnectionSettings(PhysiK::WorldHandle world, PhysiK::ComponentHandle h,
                                                 float* outStiffness, float* outDamping)
{
    const Component* c = sphere_of(W(world), h);
    if (!c)
        return;
    if (outStiffness)
        *outStiffness = c->stiffness;
    if (outDamping)
        *outDamping = c->damping;
}

// PhysiKAPI.cpp:1376-1383
int PHYSIK_GetCollisionSphereOverlapCount(PhysiK::WorldHandle world, PhysiK::ComponentHandle h)
{
    return static_cast<int>(query_overlaps(W(world), h).size());
}

// PhysiKAPI.cpp:1385-1405
int PHYSIK_GetCollisionSphereOverlaps(PhysiK::WorldHandle world, PhysiK::ComponentHandle h,
                                      PhysikCollisionSphereOverlap* outOverlaps, int maxOverlaps)
{
    if (!outOverlaps || maxOverlaps <= 0)
        return 0;
    const std::vector<PhysikCollisionSphereOverlap> o = query_overlaps(W(world), h);
    const int n = std::min(maxOverlaps, static_cast<int>(o.size()));
    std::copy(o.begin(), o.begin() + n, outOverlaps);
    return n;
}

// PhysiKAPI.cpp:1407-1427: the position is taken at the next PHYSIK_Step
void PHYSIK_SetCollisionComponentKinematicTarget(PhysiK::WorldHandle world, PhysiK::ComponentHandle h, float x,
                                                 float y, float z)
{
    Component* c = sphere_of(W(world), h);
    if (!c)
        return;
    c->target[0] = x;
    c->target[1] = y;
    c->target[2] = z;
    c->hasTarget = true;
}

int PHYSIK_GetPointConnectionCount(PhysiK::WorldHandle world)
{
    World* w = W(world);
    return w ? static_cast<int>(w->conns.size()) : 0;
}

// PhysiKAPI.cpp:1441-1475 -> World::AddPointConnection (World.cpp:126-131):
// kept only when all four nodes exist and are active.
void PHYSIK_AddPointConnection(PhysiK::WorldHandle world, int node0, int node1, int node2, int node3,
                               float barycentricX, float barycentricY, float barycentricZ, float barycentricW,
                               float targetX, float targetY, float targetZ, float stiffness, float damping)
{
    World* w = W(world);
    if (!w)
        return;
    const int n[4] = {node0, node1, node2, node3};
    for (int a = 0; a < 4; ++a)
        if (!valid_node(w, n[a]) || !w->nodeActive[n[a]])
            return;
    World::Conn c;
    std::copy(n, n + 4, c.n);
    c.w[0] = barycentricX;
    c.w[1] = barycentricY;
    c.w[2] = barycentricZ;
    c.w[3] = barycentricW;
    c.target[0] = targetX;
    c.target[1] = targetY;
    c.target[2] = targetZ;
    c.k = stiffness;
    c.c = damping;
    w->conns.push_back(c);
}

void PHYSIK_GetNodePosition(PhysiK::WorldHandle world, int i, float* outX, float* outY, float* outZ)
{
    World* w = W(world);
    if (!valid_node(w, i))
        return;
    if (const float* f = w->frame(PK_EXPORT_POSITIONS))
    {
        if (outX) *outX = f[3 * i];
        if (outY) *outY = f[3 * i + 1];
        if (outZ) *outZ = f[3 * i + 2];
        return;
    }
    w->pull();
    if (outX) *outX = static_cast<float>(w->x[3 * i]);
    if (outY) *outY = static_cast<float>(w->x[3 * i + 1]);
    if (outZ) *outZ = static_cast<float>(w->x[3 * i + 2]);
}

void PHYSIK_GetNodeVelocity(PhysiK::WorldHandle world, int i, float* outX, float* outY, float* outZ)
{
    World* w = W(world);
    if (!valid_node(w, i))
        return;
    if (const float* f = w->frame(PK_EXPORT_VELOCITIES))
    {
        if (outX) *outX = f[3 * i];
        if (outY) *outY = f[3 * i + 1];
        if (outZ) *outZ = f[3 * i + 2];
        return;
    }
    w->pull();
    if (outX) *outX = static_cast<float>(w->v[3 * i]);
    if (outY) *outY = static_cast<float>(w->v[3 * i + 1]);
    if (outZ) *outZ = static_cast<float>(w->v[3 * i + 2]);
}

void PHYSIK_SetNodePosition(PhysiK::WorldHandle world, int i, float x, float y, float z)  // World.cpp:394-401
{
    World* w = W(world);
    if (!valid_node(w, i))
        return;
    w->touch_host();
    w->x[3 * i] = x;
    w->x[3 * i + 1] = y;
    w->x[3 * i + 2] = z;
    for (int k = 0; k < 3; ++k)
        w->v[3 * i + k] = 0.0;
}

void PHYSIK_SetNodeVelocity(PhysiK::WorldHandle world, int i, float x, float y, float z)
{
    World* w = W(world);
    if (!valid_node(w, i))
        return;
    w->touch_host();
    w->v[3 * i] = x;
    w->v[3 * i + 1] = y;
    w->v[3 * i + 2] = z;
}

int PHYSIK_AMD_SetPrecision(PhysiK::WorldHandle world, int precision)
{
    World* w = W(world);
    if (!w || precision < PK_FP64 || precision > PK_FP64_UPPER)
        return 0;
    if (w->precision != precision)
        w->topoDirty = true;
    w->precision = precision;
    return 1;
}

int PHYSIK_AMD_GetLastStepStatus(PhysiK::WorldHandle world)
{
    World* w = W(world);
    return w ? w->lastStatus : PK_GROUP_NO_DYNAMIC;
}

int PHYSIK_AMD_CopyNodeStateF64(PhysiK::WorldHandle world, double* outPositions, double* outVelocities, int count)
{
    World* w = W(world);
    if (!w)
        return 0;
    w->pull();
    const int n = std::min(count, w->nodeCount());
    if (n <= 0)
        return 0;
    if (outPositions)
        std::memcpy(outPositions, w->x.data(), sizeof(double) * 3 * n);
    if (outVelocities)
        std::memcpy(outVelocities, w->v.data(), sizeof(double) * 3 * n);
    return n;
}

int PHYSIK_AMD_SetNodeStateF64(PhysiK::WorldHandle world, const double* positions, const double* velocities,
                               int count)
{
    World* w = W(world);
    if (!w)
        return 0;
    w->touch_host();
    const int n = std::min(count, w->nodeCount());
    if (n <= 0)
        return 0;
    if (positions)
        std::memcpy(w->x.data(), positions, sizeof(double) * 3 * n);
    if (velocities)
        std::memcpy(w->v.data(), velocities, sizeof(double) * 3 * n);
    return n;
}

int PHYSIK_AMD_SetFrameExport(PhysiK::WorldHandle world, int mask)
{
    World* w = W(world);
    if (!w || mask < 0 || mask > (PK_EXPORT_POSITIONS | PK_EXPORT_VELOCITIES))
        return 0;
    w->exportMask = mask;
    return 1;
}

static_assert(sizeof(PhysiK::Vec3) == 3 * sizeof(float), "Vec3 is three packed floats");

namespace
{
// Nodes [first, first + count) of positions (which = PK_EXPORT_POSITIONS) or
// velocities as float triples: the frame export while the device holds the
// newest state, else the host mirror cast to float.
int copy_node_floats(World* w, int which, int first, int count, PhysiK::Vec3* out)
{
    if (!w || !out || first < 0 || count <= 0 || first >= w->nodeCount())
        return 0;
    const int n = std::min(count, w->nodeCount() - first);
    float* dst = &out[0].x;
    if (const float* f = w->frame(which))
    {
        std::memcpy(dst, f + 3 * static_cast<size_t>(first), sizeof(float) * 3 * static_cast<size_t>(n));
        return n;
    }
    w->pull();
    const std::vector<double>& src = which == PK_EXPORT_POSITIONS ? w->x : w->v;
    for (size_t k = 0; k < 3 * static_cast<size_t>(n); ++k)
        dst[k] = static_cast<float>(src[3 * static_cast<size_t>(first) + k]);
    return n;
}
}  // namespace

int PHYSIK_AMD_GetNodePositions(PhysiK::WorldHandle world, int firstNode, int count, PhysiK::Vec3* outPositions)
{
    return copy_node_floats(W(world), PK_EXPORT_POSITIONS, firstNode, count, outPositions);
}

int PHYSIK_AMD_GetNodeVelocities(PhysiK::WorldHandle world, int firstNode, int count, PhysiK::Vec3* outVelocities)
{
    return copy_node_floats(W(world), PK_EXPORT_VELOCITIES, firstNode, count, outVelocities);
}

const char* PHYSIK_AMD_GetLastError() { return g_apiError.c_str(); }

}  // extern "C"
