// Operation counts of the oracle's substep (test and measurement infrastructure only): the oracle is
// instantiated on a counting scalar type and stepped phase by phase, so the per-unit flop constants the
// benchmark's VALU figure uses are counted from the method itself, not estimated.
//
// Convention: add, sub, mul, div, sqrt and each transcendental (log, sin, cos, acos, atan2) count one flop;
// the oracle evaluates unfused, so a multiply-add counts two. Comparisons, min / max and sign flips are
// counted apart (`cmp`) and are not flops.
//
// Scenes (config-2 material: density 0.117, EA 11473, EI 4.48e-3, h 1/6000, 8 iterations, gravity, drag):
// - solver: 64 hanging ropes x 65 vertices, nothing to touch;
// - planes: the same above one plane (z = 0) they do not reach;
// - pairs: the same in contact groups of 4, ropes far apart (every pair tested, none within reach);
// - mesh: the same with one sphere grid per rope whose lattice holds the rope, surface well clear.
// Printed as `key = value` lines.
#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

struct Ops {
    unsigned long long flop = 0, cmp = 0;
};
static Ops g_ops;

struct Cnt {
    double v;
    Cnt() : v(0) {}
    Cnt(double x) : v(x) {}
    explicit operator double() const { return v; }
    explicit operator float() const { return (float)v; }
    explicit operator int() const { return (int)v; }
    Cnt& operator+=(Cnt o) { g_ops.flop++; v += o.v; return *this; }
    Cnt& operator-=(Cnt o) { g_ops.flop++; v -= o.v; return *this; }
    Cnt& operator*=(Cnt o) { g_ops.flop++; v *= o.v; return *this; }
    Cnt& operator/=(Cnt o) { g_ops.flop++; v /= o.v; return *this; }
};
inline Cnt operator+(Cnt a, Cnt b) { g_ops.flop++; return a.v + b.v; }
inline Cnt operator-(Cnt a, Cnt b) { g_ops.flop++; return a.v - b.v; }
inline Cnt operator*(Cnt a, Cnt b) { g_ops.flop++; return a.v * b.v; }
inline Cnt operator/(Cnt a, Cnt b) { g_ops.flop++; return a.v / b.v; }
inline Cnt operator-(Cnt a) { return -a.v; }
inline bool operator<(Cnt a, Cnt b) { g_ops.cmp++; return a.v < b.v; }
inline bool operator>(Cnt a, Cnt b) { g_ops.cmp++; return a.v > b.v; }
inline bool operator<=(Cnt a, Cnt b) { g_ops.cmp++; return a.v <= b.v; }
inline bool operator>=(Cnt a, Cnt b) { g_ops.cmp++; return a.v >= b.v; }
inline bool operator==(Cnt a, Cnt b) { g_ops.cmp++; return a.v == b.v; }
inline bool operator!=(Cnt a, Cnt b) { g_ops.cmp++; return a.v != b.v; }

namespace std {
template <> class numeric_limits<Cnt> {
public:
    static constexpr bool is_specialized = true;
    static Cnt infinity() { return Cnt(numeric_limits<double>::infinity()); }
    static Cnt max() { return Cnt(numeric_limits<double>::max()); }
    static Cnt lowest() { return Cnt(numeric_limits<double>::lowest()); }
    static Cnt epsilon() { return Cnt(numeric_limits<double>::epsilon()); }
};
inline Cnt sqrt(Cnt a) { g_ops.flop++; return std::sqrt(a.v); }
inline Cnt log(Cnt a) { g_ops.flop++; return std::log(a.v); }
inline Cnt sin(Cnt a) { g_ops.flop++; return std::sin(a.v); }
inline Cnt cos(Cnt a) { g_ops.flop++; return std::cos(a.v); }
inline Cnt acos(Cnt a) { g_ops.flop++; return std::acos(a.v); }
inline Cnt atan2(Cnt a, Cnt b) { g_ops.flop++; return std::atan2(a.v, b.v); }
inline Cnt abs(Cnt a) { g_ops.cmp++; return std::abs(a.v); }
inline bool isfinite(Cnt a) { return std::isfinite(a.v); }
}  // namespace std

#include "rope_oracle.hpp"

using O = orc::Oracle<Cnt>;

struct Phases {
    unsigned long long init = 0, detect = 0, pos = 0, orient = 0, end = 0, poses = 0;
    double candidates = 0;  // vertex-mesh candidates visited per iteration, summed over substeps
};

static unsigned long long tick() { return g_ops.flop; }

// One frame of S substeps, phase by phase (the order of Oracle::substep).
static Phases frame(O& o, int S, double dt) {
    Phases p;
    const Cnt hs = Cnt(dt) / Cnt(S);
    o.begin_step();
    o.prepare(hs);
    for (int s = 0; s < S; s++) {
        unsigned long long t = tick();
        o.substep_poses(s, S);
        p.poses += tick() - t;
        t = tick();
        for (int i = 0; i < o.nv; i++) o.init_vertex(i);
        p.init += tick() - t;
        if (o.ropeContact) {
            t = tick();
            for (int i = 0; i < o.nv; i++) o.detect(i, true);
            p.detect += tick() - t;
        }
        if (!o.grids.empty())
            for (int i = 0; i < o.nv; i++)
                if (o.invMass[i] != 0) p.candidates += o.cand_count(i);
        for (int it = 0; it < o.numItr; it++) {
            t = tick();
            o.position_sweep();
            p.pos += tick() - t;
            t = tick();
            o.orientation_sweep();
            p.orient += tick() - t;
        }
        t = tick();
        for (int i = 0; i < o.nv; i++) o.end_vertex(i);
        p.end += tick() - t;
    }
    o.finish(Cnt(dt));
    o.end_step(S);
    return p;
}

// 64 ropes x 65 vertices hanging from roots on a grid `spacing` apart at height z0, 2 cm segments.
static void build(O& o, double spacing, double z0, int nropes = 64, int V = 65) {
    std::vector<float> xyz;
    std::vector<int> off{0};
    std::vector<uint8_t> pin;
    for (int r = 0; r < nropes; r++) {
        const double rx = spacing * (r % 8), ry = spacing * (r / 8);
        const double ang = 0.3 + 0.05 * r, dx = 0.3 * std::cos(ang), dy = 0.3 * std::sin(ang);
        const double n = std::sqrt(dx * dx + dy * dy + 1.0);
        for (int i = 0; i < V; i++) {
            xyz.push_back((float)(rx + 0.02 * i * dx / n));
            xyz.push_back((float)(ry + 0.02 * i * dy / n));
            xyz.push_back((float)(z0 - 0.02 * i / n));
            pin.push_back(i == 0 ? 1 : 0);
        }
        off.push_back(off.back() + V);
    }
    o.build(nropes, off.data(), xyz.data(), pin.data(), Cnt(0.117), Cnt(11473.0), Cnt(4.48e-3), Cnt(1.0 / 6000),
            8, false, false);
    o.gravity = orc::mk(Cnt(0), Cnt(0), Cnt(-9.81));
    o.drag = Cnt(0.2);
    o.damping = Cnt(0);
}

static double units(const O& o, int S) { return (double)o.nv * S; }

int main() {
    const int S = 20;
    const double dt = S / 6000.0;
    g_ops = Ops{};

    O solver;
    build(solver, 2.0, 2.0);
    frame(solver, S, dt);  // settle the first-frame special cases (lastH)
    const Phases ps = frame(solver, S, dt);
    const double U = units(solver, S), I = solver.numItr;
    const double fInit = ps.init / U, fPos = ps.pos / (U * I), fOri = ps.orient / (U * I), fEnd = ps.end / U;
    std::printf("vertices = %d\nsubsteps = %d\niterations = %d\n", solver.nv, S, solver.numItr);
    std::printf("init_flop_per_vertex_substep = %.3f\n", fInit);
    std::printf("position_sweep_flop_per_vertex_iteration = %.3f\n", fPos);
    std::printf("orientation_sweep_flop_per_vertex_iteration = %.3f\n", fOri);
    std::printf("end_flop_per_vertex_substep = %.3f\n", fEnd);
    std::printf("solver_flop_per_vertex_substep_at_8_iterations = %.3f\n", fInit + fEnd + 8 * (fPos + fOri));

    O planes;
    build(planes, 2.0, 2.0);
    planes.numPlanes = 1;
    planes.planes[0][0] = Cnt(0);
    planes.planes[0][1] = Cnt(0);
    planes.planes[0][2] = Cnt(1);
    planes.planes[0][3] = Cnt(0);
    planes.cRadius = Cnt(0.00635);
    planes.cMuS = Cnt(0.5);
    planes.cMuK = Cnt(0.4);
    frame(planes, S, dt);
    const Phases pp = frame(planes, S, dt);
    double freeV = 0;
    for (int i = 0; i < planes.nv; i++) freeV += planes.invMass[i] != 0 ? 1 : 0;
    std::printf("plane_flop_per_free_vertex_iteration = %.3f\n", (double)(pp.pos - ps.pos) / (freeV * S * I));

    O pairs;
    build(pairs, 2.0, 2.0);
    pairs.set_rope_contact(true, Cnt(0.0036), Cnt(0.0008), Cnt(0.025), Cnt(0.4), Cnt(1e7), Cnt(1.2), 4);
    frame(pairs, S, dt);
    const Phases pr = frame(pairs, S, dt);
    double tests = 0;  // pair tests of one substep (the scan's own loop bounds)
    for (int i = 0; i < pairs.nv; i++) {
        if (!(pairs.flags[i] & orc::HAS_NEXT)) continue;
        const int g = pairs.group_of(i);
        for (int j = pairs.group_begin(g); j < pairs.group_end(g); j++)
            if ((pairs.flags[j] & orc::HAS_NEXT) && !(pairs.ropeOf[j] == pairs.ropeOf[i] && std::abs(j - i) <= 2)) tests++;
    }
    std::printf("pair_tests_per_substep = %.0f\n", tests);
    std::printf("flop_per_pair_test = %.3f\n", (double)pr.detect / (tests * S));

    O mesh;
    build(mesh, 2.0, 2.0);
    {  // a sphere of radius 0.05 far below the ropes, on a lattice large enough to hold every rope of its env
        const int n = 24;
        const double h = 2.0 / (n - 1);
        std::vector<float> vals((size_t)n * n * n);
        for (int a = 0; a < n; a++)
            for (int b = 0; b < n; b++)
                for (int c = 0; c < n; c++) {
                    const double x = -1.0 + a * h, y = -1.0 + b * h, z = -1.0 + c * h;
                    vals[((size_t)a * n + b) * n + c] = (float)(std::sqrt(x * x + y * y + (z + 0.9) * (z + 0.9)) - 0.05);
                }
        std::vector<Cnt> origin{Cnt(-1.0), Cnt(-1.0), Cnt(-1.0)};
        std::vector<Cnt> p, q;
        for (int r = 0; r < mesh.nropes; r++) {
            p.push_back(Cnt(2.0 * (r % 8)));
            p.push_back(Cnt(2.0 * (r / 8)));
            p.push_back(Cnt(1.6));
            q.push_back(Cnt(0));
            q.push_back(Cnt(0));
            q.push_back(Cnt(0));
            q.push_back(Cnt(1));
        }
        mesh.add_mesh(vals.data(), n, n, n, origin.data(), Cnt(h), Cnt(0.5), Cnt(0.4), p.data(), q.data(), mesh.nropes);
    }
    frame(mesh, S, dt);
    const Phases pm = frame(mesh, S, dt);
    std::printf("mesh_candidates_per_iteration_per_frame = %.0f\n", pm.candidates);
    std::printf("flop_per_mesh_candidate_iteration = %.3f\n", (double)(pm.pos - ps.pos) / (pm.candidates * I));
    std::printf("mesh_broad_phase_flop_per_vertex_substep = %.3f\n", (double)(pm.end - ps.end) / U);
    return 0;
}
