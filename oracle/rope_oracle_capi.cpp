// C entry points over the CPU oracle (rope_oracle.hpp) for ctypes. TEST INFRASTRUCTURE ONLY:
// loaded by tests/, the smoke entry and the benchmark's CPU-baseline leg, never by the product.
// Every array crosses as float64 and is converted to the instance's precision (32 or 64).
#include <cstring>
#include <memory>
#include <string>

#include "rope_oracle.hpp"

namespace {
thread_local std::string g_err;

struct Handle {
    int prec;
    std::unique_ptr<orc::Oracle<float>> f;
    std::unique_ptr<orc::Oracle<double>> d;
};

template <class T> std::vector<T> conv(const double* p, size_t n) {
    std::vector<T> v(n);
    for (size_t i = 0; i < n; i++) v[i] = (T)p[i];
    return v;
}

template <class F> int guard(F&& fn) {
    try {
        fn();
        return 0;
    } catch (const std::invalid_argument& e) {
        g_err = e.what();
        return 1;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 3;
    }
}

template <class T> void get_impl(orc::Oracle<T>& o, const char* what, double* out) {
    const std::string w(what);
    auto put3 = [&](const std::vector<orc::V3<T>>& v) {
        for (size_t i = 0; i < v.size(); i++) {
            out[3 * i] = v[i].x;
            out[3 * i + 1] = v[i].y;
            out[3 * i + 2] = v[i].z;
        }
    };
    auto put4 = [&](const std::vector<orc::Q4<T>>& v) {
        for (size_t i = 0; i < v.size(); i++) {
            out[4 * i] = v[i].x;
            out[4 * i + 1] = v[i].y;
            out[4 * i + 2] = v[i].z;
            out[4 * i + 3] = v[i].w;
        }
    };
    auto put1 = [&](const auto& v) {
        for (size_t i = 0; i < v.size(); i++) out[i] = (double)v[i];
    };
    if (w == "positions") put3(o.pos);
    else if (w == "velocities") put3(o.vel);
    else if (w == "last_velocities") put3(o.lastVel);
    else if (w == "dx") put3(o.dx);
    else if (w == "quaternions") put4(o.q);
    else if (w == "rest_quaternions") put4(o.qRest);
    else if (w == "inv_masses") put1(o.invMass);
    else if (w == "rest_lengths") put1(o.lRest);
    else if (w == "k_stretch") put1(o.kStretch);
    else if (w == "flags") put1(o.flags);
    else if (w == "pads") put1(o.pad);
    else if (w == "max_step") put1(o.maxStep);
    else if (w == "contact_forces") put3(o.cForce);
    else if (w == "wrenches") put1(o.wrench);
    else if (w == "num_cand") put1(o.nCand);
    else if (w == "cand") put1(o.cand);
    else if (w == "num_col") put1(o.nCol);
    else if (w == "col") put1(o.coll);
    else if (w == "scalars") {
        out[0] = o.time;
        out[1] = o.h;
        out[2] = o.lastH;
        out[3] = o.radius;
        out[4] = o.barrier;
        out[5] = o.minSegLen;
        out[6] = o.lastError;
        out[7] = o.detR;
        out[8] = o.scaledDetR;
    } else
        throw std::invalid_argument("unknown field " + w);
}

template <class T> void set_scalar_impl(orc::Oracle<T>& o, const std::string& n, double v) {
    if (n == "damping") o.damping = (T)v;
    else if (n == "drag") o.drag = (T)v;
    else if (n == "drag_quad") o.dragQuad = (T)v;
    else if (n == "drag_quad_t") o.dragQuadT = (T)v;
    else if (n == "accel_ratio") o.accel = (T)v;
    else if (n == "num_itr") o.numItr = (int)v;
    else if (n == "omp_threshold") o.ompThreshold = (int)v;
    else if (n == "time") o.time = (T)v;
    else if (n == "h") o.h = (T)v;
    else throw std::invalid_argument("unknown scalar " + n);
}

template <class T> void set_state_impl(orc::Oracle<T>& o, const char* what, const double* a) {
    const std::string w(what);
    auto get3 = [&](std::vector<orc::V3<T>>& v) {
        for (size_t i = 0; i < v.size(); i++) v[i] = orc::mk((T)a[3 * i], (T)a[3 * i + 1], (T)a[3 * i + 2]);
    };
    if (w == "positions") get3(o.pos);
    else if (w == "velocities") get3(o.vel);
    else if (w == "last_velocities") get3(o.lastVel);
    else if (w == "last_positions") get3(o.lastPos);
    else if (w == "quaternions") {
        for (int i = 0; i < o.nv; i++) o.q[i] = orc::Q4<T>{(T)a[4 * i], (T)a[4 * i + 1], (T)a[4 * i + 2], (T)a[4 * i + 3]};
    } else if (w == "pads") {
        for (int i = 0; i < o.nv; i++) o.pad[i] = (T)a[i];
    } else
        throw std::invalid_argument("unknown state " + w);
}

template <class T> void planes_impl(orc::Oracle<T>& o, int n, const double* pl, double radius, double muS, double muK,
                                    double vslip, double kappa) {
    // write_planes (csrc/bindings.cpp:57-76): rows normalised, material written when n > 0
    for (int i = 0; i < n; i++) {
        orc::V3<T> nrm = orc::mk((T)pl[4 * i], (T)pl[4 * i + 1], (T)pl[4 * i + 2]);
        const T l = orc::len(nrm);
        o.planes[i][0] = nrm.x / l;
        o.planes[i][1] = nrm.y / l;
        o.planes[i][2] = nrm.z / l;
        o.planes[i][3] = (T)pl[4 * i + 3] / l;
    }
    o.numPlanes = n;
    if (n > 0) {
        o.cRadius = (T)radius;
        o.cMuS = (T)muS;
        o.cMuK = (T)muK;
        o.cVSlip = (T)vslip;
        o.cKappa = (T)kappa;
    }
}

#define DISPATCH(h, ...)                \
    do {                                \
        Handle* H_ = (Handle*)(h);      \
        if (H_->prec == 32) {           \
            auto& o = *H_->f;           \
            __VA_ARGS__;                \
        } else {                        \
            auto& o = *H_->d;           \
            __VA_ARGS__;                \
        }                               \
    } while (0)

}  // namespace

extern "C" {

const char* orc_last_error() { return g_err.c_str(); }

void* orc_create(int prec, int nropes, const int* offsets, const float* xyz, const unsigned char* pinned,
                 double density, double EA, double EI, double max_h, int num_itr, int fix_root, int straight_rest) {
    auto* H = new Handle{prec, nullptr, nullptr};
    int rc = guard([&] {
        if (prec == 32) {
            H->f.reset(new orc::Oracle<float>());
            H->f->build(nropes, offsets, xyz, pinned, (float)density, (float)EA, (float)EI, (float)max_h, num_itr,
                        fix_root != 0, straight_rest != 0);
        } else {
            H->d.reset(new orc::Oracle<double>());
            H->d->build(nropes, offsets, xyz, pinned, density, EA, EI, max_h, num_itr, fix_root != 0,
                        straight_rest != 0);
        }
    });
    if (rc) {
        delete H;
        return nullptr;
    }
    return H;
}

void orc_destroy(void* h) { delete (Handle*)h; }

int orc_num_verts(void* h) {
    int n = 0;
    DISPATCH(h, n = o.nv);
    return n;
}

int orc_get(void* h, const char* what, double* out) {
    return guard([&] { DISPATCH(h, get_impl(o, what, out)); });
}

int orc_set_state(void* h, const char* what, const double* a) {
    return guard([&] { DISPATCH(h, set_state_impl(o, what, a)); });
}

int orc_set_scalar(void* h, const char* name, double v) {
    return guard([&] { DISPATCH(h, set_scalar_impl(o, name, v)); });
}

int orc_set_gravity(void* h, double x, double y, double z) {
    return guard([&] { DISPATCH(h, o.gravity = orc::mk((decltype(o.h))x, (decltype(o.h))y, (decltype(o.h))z)); });
}

int orc_set_planes(void* h, int n, const double* planes, double radius, double muS, double muK, double vslip,
                   double kappa) {
    return guard([&] { DISPATCH(h, planes_impl(o, n, planes, radius, muS, muK, vslip, kappa)); });
}

int orc_set_contact(void* h, double radius, double muS, double muK, double vslip, double kappa) {
    return guard([&] {
        DISPATCH(h, {
            using T = decltype(o.h);
            o.cRadius = (T)radius;
            o.cMuS = (T)muS;
            o.cMuK = (T)muK;
            o.cVSlip = (T)vslip;
            o.cKappa = (T)kappa;
        });
    });
}

// which: 0 masses [nv], 1 EA per segment [nv], 2 EI per segment [nv]
int orc_set_vertex(void* h, int which, const double* a) {
    return guard([&] {
        DISPATCH(h, {
            using T = decltype(o.h);
            auto v = conv<T>(a, o.nv);
            if (which == 0) o.set_masses(v.data());
            else if (which == 1) o.set_stretch(v.data());
            else o.set_bend(v.data());
        });
    });
}

int orc_add_mesh(void* h, const float* values, int nx, int ny, int nz, const double* origin, double spacing,
                 double muS, double muK, const double* pos, const double* quat, int rows, int* index) {
    return guard([&] {
        DISPATCH(h, {
            using T = decltype(o.h);
            auto org = conv<T>(origin, 3);
            auto p = conv<T>(pos, 3 * rows);
            auto q = conv<T>(quat, 4 * rows);
            *index = o.add_mesh(values, nx, ny, nz, org.data(), (T)spacing, (T)muS, (T)muK, p.data(), q.data(), rows);
        });
    });
}

int orc_set_mesh_targets(void* h, int index, const double* pos, const double* quat, int rows,
                         const unsigned char* tele, int teleAll) {
    return guard([&] {
        DISPATCH(h, {
            using T = decltype(o.h);
            const int M = (int)o.grids.size();
            const int n = index < 0 ? o.nropes * M : rows;
            auto p = conv<T>(pos, 3 * n);
            auto q = conv<T>(quat, 4 * n);
            o.set_targets(index, p.data(), q.data(), rows, tele, teleAll != 0);
        });
    });
}

int orc_set_mesh_friction(void* h, int index, double muS, double muK) {
    return guard([&] {
        DISPATCH(h, {
            using T = decltype(o.h);
            o.grids.at(index).muS = (T)muS;
            o.grids.at(index).muK = (T)muK;
        });
    });
}

int orc_clear_meshes(void* h) { return guard([&] { DISPATCH(h, o.clear_meshes()); }); }

int orc_enable_forces(void* h, int on) {
    return guard([&] {
        DISPATCH(h, {
            o.ensure_mesh_state();
            o.forcesOn = on != 0;
        });
    });
}

int orc_set_rope_contact(void* h, int enabled, double radius, double barrier, double kc, double mu, double kf,
                         double scaler, int group) {
    return guard([&] {
        DISPATCH(h, {
            using T = decltype(o.h);
            o.set_rope_contact(enabled != 0, (T)radius, (T)barrier, (T)kc, (T)mu, (T)kf, (T)scaler, group);
        });
    });
}

int orc_step_program(void* h, const double* p0, const double* v0, const double* p1, const double* q0,
                     const double* q1, double dt, int S) {
    return guard([&] {
        DISPATCH(h, {
            using T = decltype(o.h);
            const int N = o.nropes;
            auto P0 = conv<T>(p0, 3 * N), V0 = conv<T>(v0, 3 * N), P1 = conv<T>(p1, 3 * N);
            std::vector<T> Q0, Q1;
            if (q0) Q0 = conv<T>(q0, 4 * N);
            if (q1) Q1 = conv<T>(q1, 4 * N);
            o.step_program(P0.data(), V0.data(), P1.data(), q0 ? Q0.data() : nullptr, q1 ? Q1.data() : nullptr,
                           (T)dt, S);
        });
    });
}

int orc_step_const(void* h, const double* pos, const double* vel, const double* q, double dt, int S) {
    return guard([&] {
        DISPATCH(h, {
            using T = decltype(o.h);
            const int N = o.nropes;
            auto P = conv<T>(pos, 3 * N), V = conv<T>(vel, 3 * N);
            std::vector<T> Q;
            if (q) Q = conv<T>(q, 4 * N);
            o.step_const(P.data(), V.data(), q ? Q.data() : nullptr, (T)dt, S);
        });
    });
}

int orc_step_free(void* h, double dt, int S) {
    return guard([&] { DISPATCH(h, o.step_free((decltype(o.h))dt, S)); });
}

int orc_reset_envs(void* h, const int* ids, int count) {
    return guard([&] { DISPATCH(h, o.reset_envs(ids, count)); });
}

// Rope-rope detection on given substep-start positions and pads (the device's own inputs), using the
// instance's flags, ropes, groups and radii. Outputs: counts [nv], lists [nv * 128] row-major, step limit [nv].
int orc_detect(void* h, const double* lastPos, const double* pads, int* counts, int* lists, double* maxStep) {
    return guard([&] {
        DISPATCH(h, {
            using T = decltype(o.h);
            for (int i = 0; i < o.nv; i++) {
                o.lastPos[i] = orc::mk((T)lastPos[3 * i], (T)lastPos[3 * i + 1], (T)lastPos[3 * i + 2]);
                o.pad[i] = (T)pads[i];
            }
            o.detR = o.radius + T(0.5) * o.barrier;
            o.scaledDetR = o.detR * o.detScaler;
            for (int i = 0; i < o.nv; i++) o.detect(i, true);
            for (int i = 0; i < o.nv; i++) {
                counts[i] = o.nCol[i];
                maxStep[i] = (double)o.maxStep[i];
                for (int c = 0; c < orc::MAX_COLL; c++) lists[(size_t)i * orc::MAX_COLL + c] = c < o.nCol[i] ? o.coll[(size_t)i * orc::MAX_COLL + c] : -1;
            }
        });
    });
}

// Mesh broad phase on given end-of-substep positions and displacements with the instance's current poses.
// Outputs: counts [nv] (-1 = overflow, 0 for pinned), candidate mesh ids [nv * 8] (-1 padded).
int orc_broadphase(void* h, const double* pos, const double* dx, int* counts, int* cand) {
    return guard([&] {
        DISPATCH(h, {
            using T = decltype(o.h);
            o.ensure_mesh_state();
            for (int i = 0; i < o.nv; i++) {
                orc::V3<T> p = orc::mk((T)pos[3 * i], (T)pos[3 * i + 1], (T)pos[3 * i + 2]);
                orc::V3<T> d = orc::mk((T)dx[3 * i], (T)dx[3 * i + 1], (T)dx[3 * i + 2]);
                if (o.invMass[i] != 0)
                    o.build_candidates(i, p, T(2) * orc::len(d) + T(0.02));
                else
                    o.nCand[i] = 0;
                counts[i] = o.nCand[i];
                for (int c = 0; c < orc::MAX_CAND; c++)
                    cand[(size_t)i * orc::MAX_CAND + c] =
                        (o.nCand[i] > 0 && c < o.nCand[i]) ? o.cand[(size_t)i * orc::MAX_CAND + c] : -1;
            }
        });
    });
}

// Poses the kernels would read: [nropes * 64 * 19] (pos 3, q 4, vel 3, omega 3, lo 3, hi 3).
int orc_get_poses(void* h, double* out) {
    return guard([&] {
        DISPATCH(h, {
            o.ensure_mesh_state();
            for (size_t s = 0; s < o.poses.size(); s++) {
                const auto& P = o.poses[s];
                double* w = out + s * 19;
                double v[19] = {P.pos.x, P.pos.y, P.pos.z, P.q.x, P.q.y, P.q.z, P.q.w, P.vel.x, P.vel.y, P.vel.z,
                                P.omega.x, P.omega.y, P.omega.z, P.lo.x, P.lo.y, P.lo.z, P.hi.x, P.hi.y, P.hi.z};
                std::memcpy(w, v, sizeof(v));
            }
        });
    });
}

}  // extern "C"
