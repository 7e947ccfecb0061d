// pybind11 module ropesim._ropesim: the reference's Python surface (RopeSim, RopeBatch, CudaView;
// csrc/bindings.cpp:629-1211 of the reference) over the C ABI of librope_hip.so
// (include/ropesim_hip.h). Host C++ only: validation, __cuda_array_interface__ parsing, GIL release;
// all device work happens behind the C ABI.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "ropesim_hip.h"

namespace py = pybind11;

namespace {

using farray = py::array_t<float, py::array::c_style | py::array::forcecast>;

// Status -> Python exception (ValueError for invalid input, RuntimeError otherwise).
void check(int rc) {
    if (rc == RS_OK) return;
    const std::string msg = rs_last_error_message();
    if (rc == RS_INVALID) throw py::value_error(msg);
    throw std::runtime_error(msg);
}

std::vector<float> vec3(const farray& a, const char* name) {
    if (a.ndim() != 1 || a.shape(0) != 3) throw py::value_error(std::string(name) + " must be a length-3 array");
    return {a.data()[0], a.data()[1], a.data()[2]};
}

// ---- device arrays handed in through the __cuda_array_interface__ protocol (versions 2 and 3): a dict with
// "data" = (address, read-only flag), "shape", "typestr" and "strides" in bytes (missing or None: C order).
// The reference accepts float32 arrays of an exact shape in C order (csrc/bindings.cpp:587-617), and teleport
// flags as one byte per environment (:128-172); wrong kind, shape, type or layout is a ValueError.
struct DeviceArrayIn {
    uintptr_t address = 0;
    std::string typestr;
    std::vector<py::ssize_t> shape;
    std::vector<py::ssize_t> strides;  // empty: C order
};

DeviceArrayIn read_device_array(const py::object& obj) {
    const py::dict cai = obj.attr("__cuda_array_interface__").cast<py::dict>();
    DeviceArrayIn a;
    a.address = cai["data"].cast<py::tuple>()[0].cast<uintptr_t>();
    a.typestr = cai["typestr"].cast<std::string>();
    a.shape = cai["shape"].cast<std::vector<py::ssize_t>>();
    if (cai.contains("strides") && !cai["strides"].is_none()) a.strides = cai["strides"].cast<std::vector<py::ssize_t>>();
    return a;
}

// C order for elements of `itemsize` bytes: every axis of extent > 1 steps over the product of the later extents.
bool in_c_order(const DeviceArrayIn& a, py::ssize_t itemsize) {
    if (a.strides.empty()) return true;
    if (a.strides.size() != a.shape.size()) return false;
    const size_t nd = a.shape.size();
    std::vector<py::ssize_t> packed(nd);
    py::ssize_t span = itemsize;
    for (size_t k = nd; k > 0; k--) {
        packed[k - 1] = span;
        span *= a.shape[k - 1];
    }
    for (size_t k = 0; k < nd; k++)
        if (a.shape[k] > 1 && a.strides[k] != packed[k]) return false;
    return true;
}

std::string shape_text(const std::vector<py::ssize_t>& shape) {
    std::string t = "(";
    for (size_t k = 0; k < shape.size(); k++) t += (k ? ", " : "") + std::to_string(shape[k]);
    return t + (shape.size() == 1 ? ",)" : ")");
}

// A float32 device array of exactly `want`, in C order.
const float* device_f32(const py::object& obj, const std::vector<py::ssize_t>& want, const char* name) {
    const std::string who(name);
    if (!py::hasattr(obj, "__cuda_array_interface__"))
        throw py::value_error(who + " must be a GPU array (e.g. a contiguous float32 torch tensor on cuda)");
    const DeviceArrayIn a = read_device_array(obj);
    if (a.typestr != "<f4") throw py::value_error(who + " must be float32 (got " + a.typestr + ")");
    if (a.shape != want) throw py::value_error(who + " has wrong shape " + shape_text(a.shape) + "; expected " + shape_text(want));
    if (!in_c_order(a, 4)) throw py::value_error(who + " must be C-contiguous");
    return reinterpret_cast<const float*>(a.address);
}

// Teleport flags on the device: one bool or uint8 per environment, [n], contiguous.
const uint8_t* device_flags(const py::object& obj, py::ssize_t n, const char* who) {
    const DeviceArrayIn a = read_device_array(obj);
    const bool byte = a.typestr == "|b1" || a.typestr == "|u1";
    if (!byte || a.shape != std::vector<py::ssize_t>{n} || !in_c_order(a, 1))
        throw py::value_error(std::string(who) + ": teleport must be a bool, a [num_envs] bool array, or a contiguous "
                                                 "[num_envs] GPU bool array");
    return reinterpret_cast<const uint8_t*>(a.address);
}

// Device buffer exposed to Python (torch.as_tensor(view) shares the memory).
struct CudaView {
    uintptr_t ptr;
    std::vector<py::ssize_t> shape;
};

// ---- mesh poses from the host: pos [3] with quat [4] (one pose for every environment) or [N, 3] with [N, 4]
void check_pose_rows(const float* pos, const float* quat, int rows, const char* who) {
    for (int r = 0; r < rows; r++) {
        const float* p = pos + 3 * (size_t)r;
        const float* q = quat + 4 * (size_t)r;
        double n2 = 0.0;
        for (int c = 0; c < 4; c++) n2 += double(q[c]) * double(q[c]);
        if (!std::isfinite(n2) || !(std::sqrt(n2) > 1e-6))
            throw py::value_error(std::string(who) + ": quat must be finite with nonzero norm");
        if (!(std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2])))
            throw py::value_error(std::string(who) + ": pos must be finite");
    }
}
int pose_rows(const farray& pos, const farray& quat, int N, const char* who) {
    auto rows_of = [](const farray& a, int width, int N) {
        if (a.ndim() == 1 && a.shape(0) == width) return 1;
        if (a.ndim() == 2 && a.shape(0) == N && a.shape(1) == width) return N;
        return -1;
    };
    const int rp = rows_of(pos, 3, N), rq = rows_of(quat, 4, N);
    if (rp < 0 || rp != rq)
        throw py::value_error(std::string(who) +
                              ": pos must be [3] and quat [4] (x, y, z, w), or [num_envs, 3] and [num_envs, 4]");
    check_pose_rows(pos.data(), quat.data(), rp, who);
    return rp;
}

// Owns one rs_batch.
struct Handle {
    rs_batch* b = nullptr;
    ~Handle() {
        if (b) rs_destroy(b);
    }
};

// The shared part of RopeSim and RopeBatch.
class Core {
public:
    std::shared_ptr<Handle> h;
    int N = 0, nv = 0, V = 0;  // V: vertices per rope when uniform, else 0
    std::vector<int> off;
    bool fixRoot = false;
    bool rootsPinned = true;

    rs_batch* b() const { return h->b; }
    void set_fast_math(bool on) { check(rs_set_fast_math(b(), on ? 1 : 0)); }
    bool fast_math() const {
        int on = 0;
        check(rs_get_fast_math(b(), &on));
        return on != 0;
    }

    void create(int nropes, const std::vector<int>& offsets, const float* pos, const uint8_t* pinned, float density,
                float ea, float ei, float max_h, int num_itr, bool fix_root, bool straight_rest, float min_seg) {
        rs_desc d{};
        d.num_ropes = nropes;
        d.offsets = offsets.data();
        d.positions = pos;
        d.pinned = pinned;
        d.density = density;
        d.k_stretch = ea;
        d.k_bend = ei;
        d.max_h = max_h;
        d.num_itr = num_itr;
        d.fix_root_orientation = fix_root ? 1 : 0;
        d.straight_rest = straight_rest ? 1 : 0;
        d.min_seg = min_seg;
        h = std::make_shared<Handle>();
        check(rs_create(&d, &h->b));
        N = nropes;
        off = offsets;
        nv = off[N];
        V = off[1];
        for (int r = 0; r < N; r++)
            if (off[r + 1] - off[r] != V) V = 0;
        fixRoot = fix_root;
        if (pinned)
            for (int r = 0; r < N; r++)
                if (!pinned[off[r]]) rootsPinned = false;
    }
    double param(int which) const {
        double v = 0;
        check(rs_get_param(b(), which, &v));
        return v;
    }
    void set_param(int which, double v) { check(rs_set_param(b(), which, v)); }
    std::vector<py::ssize_t> vshape(int w) const {
        if (V > 0) return {N, V, w};
        return {nv, w};
    }
    py::array_t<float> get_vertex(int which, int w) {
        py::array_t<float> out(w == 1 ? std::vector<py::ssize_t>{nv} : std::vector<py::ssize_t>{nv, w});
        check(rs_get_vertex(b(), which, out.mutable_data()));
        return out;
    }
    CudaView view(int which) {
        rs_views_t v{};
        check(rs_views(b(), &v));
        float* p = which == 0 ? v.positions : which == 1 ? v.velocities : which == 2 ? v.quaternions
                                                                                     : which == 3 ? v.contact_forces
                                                                                                  : v.wrenches;
        if (!p) throw py::value_error("enable_contact_forces(True) first");
        if (which == 4) return CudaView{(uintptr_t)p, {N, 64, 6}};
        return CudaView{(uintptr_t)p, vshape(which == 2 ? 4 : 3)};
    }
    void set_gravity(const farray& g) {
        auto v = vec3(g, "gravity");
        check(rs_set_gravity(b(), v.data()));
    }
    py::array_t<float> gravity() const {
        py::array_t<float> a(3);
        check(rs_get_gravity(b(), a.mutable_data()));
        return a;
    }
    void set_planes(const farray& planes, float radius, float mu_s, py::object mu_k, float v_slip, float kappa) {
        if (planes.ndim() != 2 || planes.shape(1) != 4 || planes.shape(0) > 8)
            throw py::value_error("planes must be [P<=8, 4] rows (nx, ny, nz, w): half-spaces n.x >= w");
        check(rs_set_planes(b(), (int)planes.shape(0), planes.data(), radius, mu_s,
                            mu_k.is_none() ? mu_s : mu_k.cast<float>(), v_slip, kappa));
    }
    void set_contact(float radius, float mu_s, py::object mu_k, float v_slip, float kappa) {
        check(rs_set_contact(b(), radius, mu_s, mu_k.is_none() ? mu_s : mu_k.cast<float>(), v_slip, kappa));
    }
    int add_mesh(const farray& values, const farray& origin, float spacing, const farray& pos, const farray& quat,
                 py::object mu_s, py::object mu_k) {
        if (values.ndim() != 3) throw py::value_error("sdf values must be [nx, ny, nz] with every dim >= 2");
        auto o = vec3(origin, "origin");
        const int rows = pose_rows(pos, quat, N, "add_mesh");
        float c[5];
        check(rs_get_contact(b(), c));
        const float ms = mu_s.is_none() ? c[1] : mu_s.cast<float>();
        const float mk = mu_k.is_none() ? (mu_s.is_none() ? c[2] : ms) : mu_k.cast<float>();
        if (!(ms >= 0) || !(mk >= 0)) throw py::value_error("mu_s / mu_k must be >= 0");
        int idx = -1;
        check(rs_add_mesh(b(), values.data(), (int)values.shape(0), (int)values.shape(1), (int)values.shape(2), o.data(),
                          spacing, ms, mk, pos.data(), quat.data(), rows, &idx));
        return idx;
    }
    void set_mesh_friction(int index, float mu_s, py::object mu_k) {
        if (!(mu_s >= 0)) throw py::value_error("mu_s must be >= 0");
        const float mk = mu_k.is_none() ? mu_s : mu_k.cast<float>();
        check(rs_set_mesh_friction(b(), index, mu_s, mk));
    }
    // teleport: bool, host [N] bool array or GPU [N] bool array
    void set_targets(int index, py::object pos, py::object quat, py::object teleport, const char* who, int M) {
        int teleAll = 0;
        const uint8_t* tele = nullptr;
        std::vector<uint8_t> hostTele;
        bool teleDev = false;
        if (py::isinstance<py::bool_>(teleport)) {
            teleAll = teleport.cast<bool>() ? 1 : 0;
        } else if (py::hasattr(teleport, "__cuda_array_interface__")) {
            tele = device_flags(teleport, N, who);
            teleDev = true;
        } else {
            auto a = py::array_t<bool, py::array::c_style | py::array::forcecast>::ensure(teleport);
            if (!a || a.ndim() != 1 || a.shape(0) != N)
                throw py::value_error(std::string(who) +
                                      ": teleport must be a bool, a [num_envs] bool array, or a [num_envs] GPU bool array");
            hostTele.resize(N);
            for (int i = 0; i < N; i++) hostTele[i] = a.data()[i] ? 1 : 0;
            tele = hostTele.data();
        }
        const bool dp = py::hasattr(pos, "__cuda_array_interface__"), dq = py::hasattr(quat, "__cuda_array_interface__");
        if (dp != dq) throw py::value_error(std::string(who) + ": pos and quat must both be host arrays or both GPU arrays");
        if (dp) {
            const int rows = index < 0 ? N * M : N;
            const float* p = index < 0 ? device_f32(pos, {N, M, 3}, "pos") : device_f32(pos, {N, 3}, "pos");
            const float* q = index < 0 ? device_f32(quat, {N, M, 4}, "quat") : device_f32(quat, {N, 4}, "quat");
            if (tele && !teleDev) {
                // host flags with device poses: stage through the host path of the flags only
                throw py::value_error(std::string(who) + ": GPU poses need teleport as a bool or a GPU bool array");
            }
            check(rs_set_mesh_targets(b(), index, p, q, index < 0 ? rows : N, tele, teleAll, 1));
            return;
        }
        if (teleDev) throw py::value_error(std::string(who) + ": host poses need teleport as a bool or a host bool array");
        farray hp = farray::ensure(pos), hq = farray::ensure(quat);
        if (!hp || !hq) throw py::value_error(std::string(who) + ": pos/quat must be arrays");
        int rows;
        if (index < 0) {
            if (hp.ndim() != 3 || hq.ndim() != 3 || hp.shape(0) != N || hq.shape(0) != N || hp.shape(1) != M ||
                hq.shape(1) != M || hp.shape(2) != 3 || hq.shape(2) != 4)
                throw py::value_error(std::string(who) +
                                      ": pos must be [num_envs, num_meshes, 3] and quat [num_envs, num_meshes, 4]");
            rows = N * M;
            check_pose_rows(hp.data(), hq.data(), rows, who);
        } else {
            rows = pose_rows(hp, hq, N, who);
        }
        check(rs_set_mesh_targets(b(), index, hp.data(), hq.data(), rows, tele, teleAll, 0));
    }
    void clear_meshes() { check(rs_clear_meshes(b())); }
    // deterministic: wrench sums in 64-bit fixed point, reproducible bit for bit (an extension; the reference's
    // float atomics are not). None leaves the current mode.
    void enable_contact_forces(bool on, py::object deterministic) {
        if (!deterministic.is_none()) check(rs_set_contact_force_mode(b(), deterministic.cast<bool>() ? 1 : 0));
        check(rs_enable_contact_forces(b(), on ? 1 : 0));
    }
    bool deterministic_forces() const {
        int on = 0;
        check(rs_get_contact_force_mode(b(), &on));
        return on != 0;
    }
    int num_meshes() const { return rs_num_meshes(b()); }
    py::array_t<float> contact_forces() {
        py::array_t<float> out(vshape(3));
        check(rs_contact_forces_host(b(), out.mutable_data()));
        return out;
    }
    py::array_t<float> mesh_wrenches() {
        const int M = num_meshes();
        py::array_t<float> out(std::vector<py::ssize_t>{N, M, 6});
        check(rs_mesh_wrenches_host(b(), out.mutable_data()));
        return out;
    }
    int last_error() const {
        int e, w;
        rs_error_codes(b(), &e, &w);
        return e;
    }
    int last_warning() const {
        int e, w;
        rs_error_codes(b(), &e, &w);
        return w;
    }
    void set_dispatch(int mode, int variant) { check(rs_set_dispatch(b(), mode, variant)); }
    // rope-rope detection of the per-substep kernels: "auto", "scan" or "bvh"
    std::string rope_broad_phase() const {
        int m = 0;
        check(rs_get_rope_broad_phase(b(), &m, nullptr));
        return m == 1 ? "scan" : m == 2 ? "bvh" : "auto";
    }
    void set_rope_broad_phase(const std::string& m) {
        if (m != "auto" && m != "scan" && m != "bvh")
            throw py::value_error("rope_broad_phase must be 'auto', 'scan' or 'bvh'");
        check(rs_set_rope_broad_phase(b(), m == "scan" ? 1 : m == "bvh" ? 2 : 0));
    }
    bool rope_bvh_active() const {
        int a = 0;
        check(rs_get_rope_broad_phase(b(), nullptr, &a));
        return a != 0;
    }
    py::tuple dispatch() const {
        int m, v;
        rs_get_dispatch(b(), &m, &v);
        return py::make_tuple(m, v);
    }
    float autotune(float dt, int substeps, bool program) {
        float ms = 0;
        py::gil_scoped_release rel;
        check(rs_autotune(b(), dt, substeps, program ? 1 : 0, &ms));
        return ms;
    }
    // Flat per-vertex array from [V] / [N, V] / [nv] (seg: [V-1] / [N, V-1] / [nv]).
    std::vector<float> per_vertex(const farray& a, bool seg, const char* what) {
        std::vector<float> out(nv, 1.f);
        const int w = seg ? -1 : 0;
        if (V > 0 && a.ndim() == 1 && a.shape(0) == V + w) {
            for (int r = 0; r < N; r++)
                for (int i = 0; i < V + w; i++) out[off[r] + i] = a.data()[i];
        } else if (V > 0 && a.ndim() == 2 && a.shape(0) == N && a.shape(1) == V + w) {
            for (int r = 0; r < N; r++)
                for (int i = 0; i < V + w; i++) out[off[r] + i] = a.data()[(size_t)r * (V + w) + i];
        } else if (a.ndim() == 1 && a.shape(0) == nv) {
            for (int i = 0; i < nv; i++) out[i] = a.data()[i];
        } else {
            throw py::value_error(std::string("expected ") + what);
        }
        return out;
    }
};

// ------------------------------------------------------------------------------------------ RopeBatch

class RopeBatch : public Core {
public:
    bool stepped = false;

    // Differentiation form (no reference counterpart): material between rollouts, recording, backward pass.
    void set_material(py::object ea, py::object ei, py::object mass) {
        const py::ssize_t n = nv;
        bool dev = false, host = false;
        for (const py::object* o : {&ea, &ei, &mass})
            if (!o->is_none()) (py::hasattr(*o, "__cuda_array_interface__") ? dev : host) = true;
        if (dev && host) throw py::value_error("set_material: pass every array on the GPU or every array on the host");
        if (dev) {
            const float* a = ea.is_none() ? nullptr : device_f32(ea, {n}, "ea");
            const float* c = ei.is_none() ? nullptr : device_f32(ei, {n}, "ei");
            const float* m = mass.is_none() ? nullptr : device_f32(mass, {n}, "mass");
            check(rs_set_material(b(), a, c, m, 1));
            return;
        }
        farray A, C, M;
        auto host_arr = [&](const py::object& o, farray& keep, const char* name) -> const float* {
            if (o.is_none()) return nullptr;
            keep = o.cast<farray>();
            if (keep.size() != n) throw py::value_error(std::string(name) + " must hold num_vertices values");
            return keep.data();
        };
        const float* a = host_arr(ea, A, "ea");
        const float* c = host_arr(ei, C, "ei");
        const float* m = host_arr(mass, M, "mass");
        check(rs_set_material(b(), a, c, m, 0));
    }
    void set_graph(bool on) { check(rs_set_graph(b(), on ? 1 : 0)); }
    long long graph_captures() const {
        long long c = 0;
        check(rs_graph_captures(b(), &c));
        return c;
    }
    void tape_begin(int max_frames) { check(rs_tape_begin(b(), max_frames)); }
    void tape_end() { check(rs_tape_end(b())); }
    int tape_frames() const { return rs_tape_frames(b()); }
    py::tuple adjoint(py::object dl_dx, py::object g_ea, py::object g_ei, py::object g_mass, py::object g_p0,
                      py::object g_v0, py::object g_p1) {
        const py::ssize_t F = rs_tape_frames(b()), n = nv, R = N;
        const float* dl = device_f32(dl_dx, {F, n, 3}, "dl_dx");
        auto out = [&](const py::object& o, const std::vector<py::ssize_t>& shape, const char* name) -> float* {
            return o.is_none() ? nullptr : const_cast<float*>(device_f32(o, shape, name));
        };
        float* ge = out(g_ea, {n}, "g_ea");
        float* gi = out(g_ei, {n}, "g_ei");
        float* gm = out(g_mass, {n}, "g_mass");
        float* gp0 = out(g_p0, {F, R, 3}, "g_p0");
        float* gv0 = out(g_v0, {F, R, 3}, "g_v0");
        float* gp1 = out(g_p1, {F, R, 3}, "g_p1");
        double gd = 0.0;
        int mis = 0, rc;
        {
            py::gil_scoped_release rel;
            rc = rs_adjoint(b(), dl, ge, gi, gm, &gd, gp0, gv0, gp1, &mis);
        }
        check(rc);
        return py::make_tuple(gd, mis);
    }

    void step(py::object root_pos, py::object root_vel, float dt, int substeps, py::object root_q,
              py::object root_pos_end, py::object root_q_end) {
        if (!(dt > 0)) throw py::value_error("dt must be > 0");
        if (substeps < 1) throw py::value_error("substeps must be >= 1");
        if (!rootsPinned)
            throw py::value_error("step() drives every rope's root; ropes built with a free root "
                                  "(from_positions pinned=...) advance with step_free()");
        const float* p = device_f32(root_pos, {N, 3}, "root_pos");
        const float* v = device_f32(root_vel, {N, 3}, "root_vel");
        const float* q = nullptr;
        if (!root_q.is_none()) {
            if (!fixRoot) throw py::value_error("root_q requires fix_root_orientation=True");
            q = device_f32(root_q, {N, 4}, "root_q");
        }
        if (root_pos_end.is_none()) {
            if (!root_q_end.is_none()) throw py::value_error("root_q_end requires root_pos_end (the motion-program path)");
            stepped = true;
            int rc;
            {
                py::gil_scoped_release rel;
                rc = rs_step_const(b(), p, v, q, dt, substeps);
            }
            check(rc);
            return;
        }
        const float* p1 = device_f32(root_pos_end, {N, 3}, "root_pos_end");
        const float* q1 = nullptr;
        if (!root_q_end.is_none()) {
            if (!q) throw py::value_error("root_q_end requires root_q");
            q1 = device_f32(root_q_end, {N, 4}, "root_q_end");
        }
        stepped = true;
        int rc;
        {
            py::gil_scoped_release rel;
            rc = rs_step_program(b(), p, v, p1, q, q1, dt, substeps);
        }
        check(rc);
    }
    void step_free(float dt, int substeps) {
        if (!(dt > 0) || substeps < 1) throw py::value_error("dt > 0 and substeps >= 1 required");
        stepped = true;
        int rc;
        {
            py::gil_scoped_release rel;
            rc = rs_step_free(b(), dt, substeps);
        }
        check(rc);
    }
    void reset_envs(py::array_t<int, py::array::c_style | py::array::forcecast> ids) {
        if (ids.ndim() != 1) throw py::value_error("ids must be 1-D");
        if (ids.shape(0) > N) throw py::value_error("ids longer than num_envs");
        std::vector<int> v(ids.data(), ids.data() + ids.shape(0));
        for (int i : v)
            if (i < 0 || i >= N) throw py::value_error("env id out of range");
        int rc;
        {
            py::gil_scoped_release rel;
            rc = rs_reset_envs(b(), v.data(), (int)v.size());
        }
        check(rc);
    }
    void set_velocities(const farray& v) {
        const bool ok = (V > 0 && v.ndim() == 3 && v.shape(0) == N && v.shape(1) == V && v.shape(2) == 3) ||
                        (v.ndim() == 2 && v.shape(0) == nv && v.shape(1) == 3);
        if (!ok) throw py::value_error(V > 0 ? "expected [num_envs, V, 3]" : "expected [num_vertices, 3]");
        check(rs_set_vertex(b(), RS_V_VELOCITIES, v.data()));
    }
    void set_rope_contact(bool enabled, float radius, float barrier, float kc, float mu, float kf, float scaler,
                          int group) {
        check(rs_set_rope_contact(b(), enabled ? 1 : 0, radius, barrier, kc, mu, kf, scaler, group));
    }
    py::array_t<float> inv_masses() {
        auto a = get_vertex(RS_V_INV_MASSES, 1);
        if (V > 0) a.resize({N, V});
        return a;
    }
    void set_per_vertex(int which, const farray& a, const char* name, const char* shapes) {
        if (stepped) throw py::value_error(std::string(name) + " is only legal before the first step");
        auto flat = per_vertex(a, which != RS_V_MASSES, shapes);
        check(rs_set_vertex(b(), which, flat.data()));
    }
};

std::unique_ptr<RopeBatch> make_batch_uniform(int num_envs, int V, const farray& root_pos, float seg_len,
                                              const farray& dir, float density, float k_stretch, float k_bend,
                                              float max_h, int num_itr, bool fix_root) {
    if (num_envs < 1) throw py::value_error("num_envs must be >= 1");
    if (V < 3) throw py::value_error("verts_per_rope must be >= 3");
    if (!(seg_len > 0)) throw py::value_error("seg_len must be > 0");
    if (!(density > 0)) throw py::value_error("density must be > 0");
    if (!(k_stretch > 0)) throw py::value_error("k_stretch must be > 0");
    if (!(k_bend > 0)) throw py::value_error("k_bend must be > 0 (0 -> NaN in the frame update)");
    if (!(max_h > 0)) throw py::value_error("max_h must be > 0");
    if (num_itr < 1) throw py::value_error("num_itr must be >= 1");
    if (root_pos.ndim() != 2 || root_pos.shape(0) != num_envs || root_pos.shape(1) != 3)
        throw py::value_error("root_pos must be a host [num_envs, 3] array");
    auto d = vec3(dir, "dir");
    const float dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(dl > 0)) throw py::value_error("dir must be nonzero");
    const float inv = 1.f / dl;  // normalize (csrc/batch.cu:206)
    const float dn[3] = {d[0] * inv, d[1] * inv, d[2] * inv};
    std::vector<float> pos((size_t)num_envs * V * 3);
    for (int e = 0; e < num_envs; e++)
        for (int v = 0; v < V; v++)
            for (int c = 0; c < 3; c++)
                pos[((size_t)e * V + v) * 3 + c] = root_pos.data()[3 * e + c] + (float(v) * seg_len) * dn[c];
    std::vector<int> off(num_envs + 1);
    for (int e = 0; e <= num_envs; e++) off[e] = e * V;
    auto b = std::make_unique<RopeBatch>();
    b->create(num_envs, off, pos.data(), nullptr, density, k_stretch, k_bend, max_h, num_itr, fix_root, false, seg_len);
    return b;
}

// positions: [N, V, 3], a list of [V_i, 3] arrays, or flat [sum V, 3] with offsets.
std::unique_ptr<RopeBatch> make_batch_positions(py::object positions, float density, float k_stretch, float k_bend,
                                                float max_h, int num_itr, bool fix_root, py::object pinned,
                                                bool straight_rest, py::object offsets) {
    std::vector<float> flat;
    std::vector<int> off;
    int V = 0;
    if (!offsets.is_none()) {
        farray p = farray::ensure(positions);
        auto o = py::array_t<int, py::array::c_style | py::array::forcecast>::ensure(offsets);
        if (!p || p.ndim() != 2 || p.shape(1) != 3) throw py::value_error("flat positions must be [num_vertices, 3]");
        if (!o || o.ndim() != 1 || o.shape(0) < 2) throw py::value_error("offsets must be [num_ropes + 1]");
        off.assign(o.data(), o.data() + o.shape(0));
        if (off.back() != p.shape(0)) throw py::value_error("offsets[-1] must equal the number of vertices");
        flat.assign(p.data(), p.data() + p.size());
    } else if (py::isinstance<py::list>(positions) || py::isinstance<py::tuple>(positions)) {
        off.push_back(0);
        for (auto item : positions) {
            farray r = farray::ensure(item);
            if (!r || r.ndim() != 2 || r.shape(1) != 3 || r.shape(0) < 3) throw py::value_error("each rope must be [V_i >= 3, 3]");
            flat.insert(flat.end(), r.data(), r.data() + r.size());
            off.push_back(off.back() + (int)r.shape(0));
        }
    } else {
        farray p = farray::ensure(positions);
        if (!p || p.ndim() != 3 || p.shape(2) != 3 || p.shape(1) < 3) throw py::value_error("positions must be [num_envs, V>=3, 3]");
        const int N = (int)p.shape(0);
        V = (int)p.shape(1);
        for (int e = 0; e <= N; e++) off.push_back(e * V);
        flat.assign(p.data(), p.data() + p.size());
    }
    const int N = (int)off.size() - 1;
    if (N < 1) throw py::value_error("need at least one rope");
    for (int r = 0; r < N; r++)
        if (off[r + 1] - off[r] < 3) throw py::value_error("every rope needs >= 3 vertices");
    if (!(density > 0) || !(k_stretch > 0) || !(k_bend > 0) || !(max_h > 0) || num_itr < 1)
        throw py::value_error("density, k_stretch, k_bend, max_h > 0 and num_itr >= 1 required");
    const int nv = off[N];
    std::vector<uint8_t> pin(nv, 0);
    if (pinned.is_none()) {
        for (int r = 0; r < N; r++) pin[off[r]] = 1;
    } else {
        auto pm = py::array_t<bool, py::array::c_style | py::array::forcecast>::ensure(pinned);
        if (!pm) throw py::value_error("pinned must be a bool array");
        int Vu = off[1];
        for (int r = 0; r < N; r++)
            if (off[r + 1] - off[r] != Vu) Vu = 0;
        if (Vu > 0 && pm.ndim() == 1 && pm.shape(0) == Vu) {
            for (int r = 0; r < N; r++)
                for (int v = 0; v < Vu; v++) pin[off[r] + v] = pm.data()[v];
        } else if (Vu > 0 && pm.ndim() == 2 && pm.shape(0) == N && pm.shape(1) == Vu) {
            for (int i = 0; i < nv; i++) pin[i] = pm.data()[i];
        } else if (pm.ndim() == 1 && pm.shape(0) == nv) {
            for (int i = 0; i < nv; i++) pin[i] = pm.data()[i];
        } else {
            throw py::value_error("pinned must be [V] or [num_envs, V] booleans (flat [num_vertices] for uneven ropes)");
        }
    }
    if (fix_root)
        for (int r = 0; r < N; r++)
            if (!pin[off[r]]) throw py::value_error("fix_root_orientation=True needs every root pinned");
    for (float x : flat)
        if (!std::isfinite(x)) throw py::value_error("positions must be finite");
    auto b = std::make_unique<RopeBatch>();
    b->create(N, off, flat.data(), pin.data(), density, k_stretch, k_bend, max_h, num_itr, fix_root, straight_rest, 0.f);
    return b;
}

// ------------------------------------------------------------------------------------------ RopeSim

class RopeSim : public Core {
public:
    int n = 0;
    bool collisions = false;

    RopeSim(int num_verts, float seg_len, const farray& dir, float density, float k_stretch, float k_bend, float max_h,
            int num_itr, bool fix_root_orientation, bool collisions_) {
        if (num_verts < 3) throw py::value_error("num_verts must be >= 3");
        if (!(seg_len > 0)) throw py::value_error("seg_len must be > 0");
        if (!(density > 0)) throw py::value_error("density must be > 0");
        if (!(k_stretch > 0)) throw py::value_error("k_stretch must be > 0");
        if (!(k_bend > 0))
            throw py::value_error("k_bend must be > 0: k_bend=0 zeroes the rest-frame vectors and the orientation solve "
                                  "produces NaNs (normalize of a zero vector)");
        if (!(max_h > 0)) throw py::value_error("max_h must be > 0");
        if (num_itr < 1) throw py::value_error("num_itr must be >= 1");
        auto d = vec3(dir, "dir");
        const float dl = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (!(dl > 0)) throw py::value_error("dir must be nonzero");
        if (collisions_ && 2 * seg_len < 2 * 1e-4f + 8e-4f)
            throw py::value_error("collisions=True uses the upstream yarn-scale collision radius (1e-4 m): seg_len must "
                                  "be >= 5e-4 m");
        n = num_verts;
        collisions = collisions_;
        std::vector<float> pos((size_t)n * 3);
        for (int i = 0; i < n; i++)
            for (int c = 0; c < 3; c++) pos[3 * i + c] = float(i) * seg_len * (d[c] / dl);
        std::vector<int> off{0, n};
        std::vector<uint8_t> pin(n, 0);
        pin[0] = 1;
        create(1, off, pos.data(), pin.data(), density, k_stretch, k_bend, max_h, num_itr, fix_root_orientation, false,
               seg_len);
        if (collisions) {
            // the upstream yarn defaults (YB/YarnBall.cu:33-43), self-contact of the one rope
            check(rs_set_rope_contact(b(), 1, 1e-4f, 8e-4f, 1e-5f, 0.1f, 5.f, 1.2f, 1));
        }
    }
    // advance(dt, substeps=0): 0 -> ceil(dt / max_h) in float32; substeps > 0 -> exactly that many of dt / substeps
    // (csrc/bindings.cpp:270-295).
    void advance(float dt, int substeps) {
        if (substeps < 0) throw py::value_error("substeps must be >= 0");
        if (!(dt > 0)) return;
        int S = substeps;
        if (S == 0) {
            const float mh = (float)param(RS_P_MAX_H);
            S = std::max(1, (int)std::ceil(dt / mh));
        }
        int rc;
        {
            py::gil_scoped_release rel;
            rc = rs_step_free(b(), dt, S);
        }
        check(rc);
    }
    void reset() {
        int id = 0;
        check(rs_reset_envs(b(), &id, 1));
        check(rs_reset_clock(b()));
    }
    void set_root(const farray& pos, const farray& vel, py::object q) {
        auto p = vec3(pos, "pos"), v = vec3(vel, "vel");
        std::vector<float> qq;
        if (!q.is_none()) {
            if (!fixRoot) throw py::value_error("set_root(q=...) requires fix_root_orientation=True");
            farray qa = q.cast<farray>();
            if (qa.ndim() != 1 || qa.shape(0) != 4) throw py::value_error("q must be a length-4 array (x,y,z,w)");
            const float* a = qa.data();
            const float l = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2] + a[3] * a[3]);
            if (!std::isfinite(l) || !(l > 1e-6f)) throw py::value_error("q must be finite with nonzero norm");
            qq = {a[0] / l, a[1] / l, a[2] / l, a[3] / l};
        }
        auto im = get_vertex(RS_V_INV_MASSES, 1);
        if (im.data()[0] != 0.f)
            throw py::value_error("set_root drives a pinned root; this rope's root was unpinned (unpin_root)");
        check(rs_set_root_host(b(), 0, p.data(), v.data(), qq.empty() ? nullptr : qq.data()));
    }
    py::array_t<float> rest_lengths() {
        auto a = get_vertex(RS_V_REST_LENGTHS, 1);
        py::array_t<float> out(n - 1);
        for (int i = 0; i < n - 1; i++) out.mutable_data()[i] = a.data()[i];
        return out;
    }
    void set_seg(int which, const farray& a, const char* name) {
        if (a.ndim() != 1 || a.shape(0) != n - 1) throw py::value_error(std::string("expected length n-1 array of ") + name + " values");
        std::vector<float> v(n, 1.f);
        for (int i = 0; i < n - 1; i++) {
            if (!std::isfinite(a.data()[i]) || !(a.data()[i] > 0))
                throw py::value_error(std::string(name) + " values must be finite and > 0");
            v[i] = a.data()[i];
        }
        check(rs_set_vertex(b(), which, v.data()));
    }
    void set_velocities(const farray& v) {
        if (v.ndim() != 2 || v.shape(0) != n || v.shape(1) != 3) throw py::value_error("expected [n,3] array of velocities");
        check(rs_set_vertex(b(), RS_V_VELOCITIES, v.data()));
    }
    void unpin_root() {  // csrc/bindings.cpp:520-526: m0 = m1 / 2 for uniform segments
        auto im = get_vertex(RS_V_INV_MASSES, 1);
        std::vector<float> a(im.data(), im.data() + n);
        a[0] = 2.f * a[1];
        check(rs_set_vertex(b(), RS_V_INV_MASSES, a.data()));
        auto vel = get_vertex(RS_V_VELOCITIES, 3);
        std::vector<float> vv(vel.data(), vel.data() + 3 * n);
        vv[0] = vv[1] = vv[2] = 0.f;
        check(rs_set_vertex(b(), RS_V_VELOCITIES, vv.data()));
    }
    py::array_t<float> contact_forces_1() {
        py::array_t<float> out(std::vector<py::ssize_t>{n, 3});
        check(rs_contact_forces_host(b(), out.mutable_data()));
        return out;
    }
    py::array_t<float> mesh_wrenches_1() {
        const int M = num_meshes();
        py::array_t<float> out(std::vector<py::ssize_t>{M, 6});
        check(rs_mesh_wrenches_host(b(), out.mutable_data()));
        return out;
    }
};

}  // namespace

PYBIND11_MODULE(_ropesim, m) {
    m.doc() = "MI355X Cosserat rope solver: RopeSim (single rope, host API), RopeBatch (device API), CudaView.";
    m.attr("ABI_VERSION") = rs_abi_version();

    py::class_<CudaView>(m, "CudaView")
        .def_property_readonly("__cuda_array_interface__",
                               [](const CudaView& v) {
                                   py::dict d;
                                   py::tuple shape(v.shape.size());
                                   for (size_t i = 0; i < v.shape.size(); i++) shape[i] = v.shape[i];
                                   d["shape"] = shape;
                                   d["typestr"] = "<f4";
                                   d["data"] = py::make_tuple(v.ptr, false);
                                   d["strides"] = py::none();
                                   d["version"] = 2;
                                   return d;
                               })
        .def_property_readonly("ptr", [](const CudaView& v) { return v.ptr; })
        .def_property_readonly("shape", [](const CudaView& v) {
            py::tuple t(v.shape.size());
            for (size_t i = 0; i < v.shape.size(); i++) t[i] = v.shape[i];
            return t;
        });

    m.def("num_cell_variants", &rs_num_cell_variants);
    // ropesim.sdf.bake_mesh's device pass (validation and the lattice are on the Python side, as in the reference)
    m.def(
        "bake_mesh_sdf",
        [](const farray& vertices, const py::array_t<int, py::array::c_style | py::array::forcecast>& triangles,
           const farray& origin, float spacing, std::vector<int> dims, float thickness, float round, int device) {
            if (vertices.ndim() != 2 || vertices.shape(1) != 3) throw py::value_error("vertices must be [V, 3]");
            if (triangles.ndim() != 2 || triangles.shape(1) != 3) throw py::value_error("triangles must be [T, 3]");
            if (origin.size() != 3 || dims.size() != 3) throw py::value_error("origin and dims need 3 entries");
            py::array_t<float> out({(py::ssize_t)dims[0], (py::ssize_t)dims[1], (py::ssize_t)dims[2]});
            float* o = out.mutable_data();
            const float* v = vertices.data();
            const int* t = triangles.data();
            const float* org = origin.data();
            const int nv = (int)vertices.shape(0), nt = (int)triangles.shape(0);
            int st;
            {
                py::gil_scoped_release nogil;
                st = rs_bake_mesh_sdf(v, nv, t, nt, org, spacing, dims[0], dims[1], dims[2], thickness, round, device, o);
            }
            check(st);
            return out;
        },
        py::arg("vertices"), py::arg("triangles"), py::arg("origin"), py::arg("spacing"), py::arg("dims"),
        py::arg("thickness") = 0.f, py::arg("round") = 0.f, py::arg("device") = -1);
    m.def("cell_variant_info", [](int v) {
        int b = 0, t = 0, s = 0;
        check(rs_cell_variant_info(v, &b, &t, &s));
        return py::make_tuple(b, t, s != 0);
    });

    py::class_<RopeSim>(m, "RopeSim")
        .def_property("fast_math", &RopeSim::fast_math, &RopeSim::set_fast_math,
             "Arithmetic of the per-iteration routines: True (default) fused multiply-add and the hardware "
             "reciprocal / square root / log (about 1 ulp); False every operation correctly rounded and unfused, "
             "as the float32 oracle")
        .def(py::init<int, float, const farray&, float, float, float, float, int, bool, bool>(), py::arg("num_verts"),
             py::arg("seg_len") = 0.02f, py::arg("dir") = farray(py::cast(std::vector<float>{1, 0, 0})),
             py::arg("density") = 0.1f, py::arg("k_stretch") = 500.0f, py::arg("k_bend") = 1e-6f,
             py::arg("max_h") = 1.f / 3000.f, py::arg("num_itr") = 8, py::arg("fix_root_orientation") = false,
             py::arg("collisions") = false)
        .def("advance", &RopeSim::advance, py::arg("dt"), py::arg("substeps") = 0)
        .def("reset", &RopeSim::reset)
        .def("set_root", &RopeSim::set_root, py::arg("pos"), py::arg("vel"), py::arg("q") = py::none())
        .def("positions", [](RopeSim& s) { return s.get_vertex(RS_V_POSITIONS, 3); })
        .def("velocities", [](RopeSim& s) { return s.get_vertex(RS_V_VELOCITIES, 3); })
        .def("quaternions", [](RopeSim& s) { return s.get_vertex(RS_V_QUATERNIONS, 4); })
        .def("inv_masses", [](RopeSim& s) { return s.get_vertex(RS_V_INV_MASSES, 1); })
        .def("rest_lengths", &RopeSim::rest_lengths)
        .def("set_stretch_stiffnesses", [](RopeSim& s, const farray& a) { s.set_seg(RS_V_EA, a, "EA"); }, py::arg("ea"))
        .def("set_bend_stiffnesses", [](RopeSim& s, const farray& a) { s.set_seg(RS_V_EI, a, "EI"); }, py::arg("ei"))
        .def("set_gravity", &RopeSim::set_gravity, py::arg("gravity"))
        .def("set_planes", &RopeSim::set_planes, py::arg("planes"), py::arg("radius") = 0.f, py::arg("mu_s") = 0.5f,
             py::arg("mu_k") = py::none(), py::arg("v_slip") = 1e-3f, py::arg("kappa") = 2.f)
        .def("set_contact", &RopeSim::set_contact, py::arg("radius"), py::arg("mu_s") = 0.5f, py::arg("mu_k") = py::none(),
             py::arg("v_slip") = 1e-3f, py::arg("kappa") = 2.f)
        .def("unpin_root", &RopeSim::unpin_root)
        .def("add_mesh", &RopeSim::add_mesh, py::arg("values"), py::arg("origin"), py::arg("spacing"),
             py::arg("pos") = farray(py::cast(std::vector<float>{0, 0, 0})),
             py::arg("quat") = farray(py::cast(std::vector<float>{0, 0, 0, 1})), py::arg("mu_s") = py::none(),
             py::arg("mu_k") = py::none())
        .def("set_mesh_friction", &RopeSim::set_mesh_friction, py::arg("index"), py::arg("mu_s"), py::arg("mu_k") = py::none())
        .def("set_mesh_pose",
             [](RopeSim& s, int index, const farray& pos, const farray& quat, bool teleport) {
                 pose_rows(pos, quat, 1, "set_mesh_pose");
                 s.set_targets(index, pos, quat, py::bool_(teleport), "set_mesh_pose", s.num_meshes());
             },
             py::arg("index"), py::arg("pos"), py::arg("quat"), py::arg("teleport") = false)
        .def("clear_meshes", &RopeSim::clear_meshes)
        .def("enable_contact_forces", &RopeSim::enable_contact_forces, py::arg("on") = true,
             py::arg("deterministic") = py::none())
        .def_property_readonly("deterministic_forces", &RopeSim::deterministic_forces)
        .def("contact_forces", &RopeSim::contact_forces_1)
        .def("mesh_wrenches", &RopeSim::mesh_wrenches_1)
        .def("set_velocities", &RopeSim::set_velocities, py::arg("velocities"))
        .def("set_dispatch", &RopeSim::set_dispatch, py::arg("mode"), py::arg("variant") = -1)
        .def_property_readonly("dispatch", &RopeSim::dispatch)
        .def_property("rope_broad_phase", &RopeSim::rope_broad_phase, &RopeSim::set_rope_broad_phase)
        .def_property_readonly("rope_bvh_active", &RopeSim::rope_bvh_active)
        .def_property("drag", [](RopeSim& s) { return s.param(RS_P_DRAG); }, [](RopeSim& s, float v) { s.set_param(RS_P_DRAG, v); })
        .def_property("drag_quad", [](RopeSim& s) { return s.param(RS_P_DRAG_QUAD); },
                      [](RopeSim& s, float v) { s.set_param(RS_P_DRAG_QUAD, v); })
        .def_property("drag_quad_t", [](RopeSim& s) { return s.param(RS_P_DRAG_QUAD_T); },
                      [](RopeSim& s, float v) { s.set_param(RS_P_DRAG_QUAD_T, v); })
        .def_property("damping", [](RopeSim& s) { return s.param(RS_P_DAMPING); },
                      [](RopeSim& s, float v) { s.set_param(RS_P_DAMPING, v); })
        .def_property("num_itr", [](RopeSim& s) { return (int)s.param(RS_P_NUM_ITR); },
                      [](RopeSim& s, int v) { s.set_param(RS_P_NUM_ITR, v); })
        .def_property("accel_ratio", [](RopeSim& s) { return s.param(RS_P_ACCEL); },
                      [](RopeSim& s, float v) { s.set_param(RS_P_ACCEL, v); })
        .def_property("max_h", [](RopeSim& s) { return s.param(RS_P_MAX_H); },
                      [](RopeSim& s, float v) { s.set_param(RS_P_MAX_H, v); })
        .def_property_readonly("gravity", &RopeSim::gravity)
        .def_property_readonly("time", [](RopeSim& s) { return s.param(RS_P_TIME); })
        .def_property_readonly("last_error", &RopeSim::last_error)
        .def_property_readonly("last_warning", &RopeSim::last_warning)
        .def_property_readonly("collisions_enabled", [](RopeSim& s) { return s.collisions; })
        .def_property_readonly("fixed_root_orientation", [](RopeSim& s) { return s.fixRoot; })
        .def_property_readonly("num_verts", [](RopeSim& s) { return s.n; });

    py::class_<RopeBatch>(m, "RopeBatch")
        .def_property("fast_math", &RopeBatch::fast_math, &RopeBatch::set_fast_math,
             "Arithmetic of the per-iteration routines: True (default) fused multiply-add and the hardware "
             "reciprocal / square root / log (about 1 ulp); False every operation correctly rounded and unfused, "
             "as the float32 oracle")
        .def(py::init([](int num_envs, int V, const farray& root_pos, float seg_len, const farray& dir, float density,
                         float k_stretch, float k_bend, float max_h, int num_itr, bool fix_root, bool /*deterministic*/) {
                 return make_batch_uniform(num_envs, V, root_pos, seg_len, dir, density, k_stretch, k_bend, max_h, num_itr,
                                           fix_root);
             }),
             py::arg("num_envs"), py::arg("verts_per_rope"), py::arg("root_pos"), py::arg("seg_len") = 0.02f,
             py::arg("dir") = farray(py::cast(std::vector<float>{0, 0, -1})), py::arg("density") = 0.1f,
             py::arg("k_stretch") = 500.0f, py::arg("k_bend") = 1e-6f, py::arg("max_h") = 1.f / 3000.f,
             py::arg("num_itr") = 8, py::arg("fix_root_orientation") = true, py::arg("deterministic") = false,
             "N straight ropes of V vertices from host roots [N, 3] along dir. Sweeps are always the deterministic "
             "(snapshot) form, so `deterministic` is accepted for compatibility and has no effect.")
        .def_static("from_positions",
                    [](py::object positions, float density, float k_stretch, float k_bend, float max_h, int num_itr,
                       bool fix_root, py::object pinned, bool straight_rest, bool /*deterministic*/, py::object offsets) {
                        return make_batch_positions(positions, density, k_stretch, k_bend, max_h, num_itr, fix_root,
                                                    pinned, straight_rest, offsets);
                    },
                    py::arg("positions"), py::arg("density"), py::arg("k_stretch"), py::arg("k_bend"), py::arg("max_h"),
                    py::arg("num_itr"), py::arg("fix_root_orientation") = false, py::arg("pinned") = py::none(),
                    py::arg("straight_rest") = false, py::arg("deterministic") = false, py::arg("offsets") = py::none(),
                    "Build ropes from vertex positions: [N, V, 3], a list of [V_i, 3] arrays, or flat [sum V, 3] with "
                    "offsets [N + 1]. Sweeps are always the deterministic (snapshot) form.")
        .def("step", &RopeBatch::step, py::arg("root_pos"), py::arg("root_vel"), py::arg("dt"), py::arg("substeps"),
             py::arg("root_q") = py::none(), py::arg("root_pos_end") = py::none(), py::arg("root_q_end") = py::none())
        .def("step_free", &RopeBatch::step_free, py::arg("dt"), py::arg("substeps"))
        .def("reset_envs", &RopeBatch::reset_envs, py::arg("ids"))
        .def("gather", [](RopeBatch& b) {
            int rc;
            {
                py::gil_scoped_release rel;
                rc = rs_gather(b.b());
            }
            check(rc);
        })
        .def("positions", [](RopeBatch& b) { return b.view(0); }, py::keep_alive<0, 1>())
        .def("velocities", [](RopeBatch& b) { return b.view(1); }, py::keep_alive<0, 1>())
        .def("quaternions", [](RopeBatch& b) { return b.view(2); }, py::keep_alive<0, 1>())
        .def("contact_forces_view", [](RopeBatch& b) { return b.view(3); }, py::keep_alive<0, 1>())
        .def("mesh_wrenches_view", [](RopeBatch& b) { return b.view(4); }, py::keep_alive<0, 1>())
        .def("contact_forces", &RopeBatch::contact_forces)
        .def("mesh_wrenches", &RopeBatch::mesh_wrenches)
        .def("set_velocities", &RopeBatch::set_velocities, py::arg("velocities"))
        .def("set_planes", &RopeBatch::set_planes, py::arg("planes"), py::arg("radius") = 0.f, py::arg("mu_s") = 0.5f,
             py::arg("mu_k") = py::none(), py::arg("v_slip") = 1e-3f, py::arg("kappa") = 2.f)
        .def("set_contact", &RopeBatch::set_contact, py::arg("radius"), py::arg("mu_s") = 0.5f, py::arg("mu_k") = py::none(),
             py::arg("v_slip") = 1e-3f, py::arg("kappa") = 2.f)
        .def("set_rope_contact", &RopeBatch::set_rope_contact, py::arg("enabled"), py::arg("radius") = 0.f,
             py::arg("barrier") = 0.f, py::arg("k_collision") = 0.f, py::arg("mu") = 0.f, py::arg("k_friction") = 0.f,
             py::arg("detection_scaler") = 1.2f, py::arg("group_ropes") = 1)
        .def("add_mesh", &RopeBatch::add_mesh, py::arg("values"), py::arg("origin"), py::arg("spacing"),
             py::arg("pos") = farray(py::cast(std::vector<float>{0, 0, 0})),
             py::arg("quat") = farray(py::cast(std::vector<float>{0, 0, 0, 1})), py::arg("mu_s") = py::none(),
             py::arg("mu_k") = py::none())
        .def("set_mesh_friction", &RopeBatch::set_mesh_friction, py::arg("index"), py::arg("mu_s"), py::arg("mu_k") = py::none())
        .def("set_mesh_poses",
             [](RopeBatch& b, int index, py::object pos, py::object quat, py::object teleport) {
                 b.set_targets(index, pos, quat, teleport, "set_mesh_poses", b.num_meshes());
             },
             py::arg("index"), py::arg("pos"), py::arg("quat"), py::arg("teleport") = false)
        .def("set_mesh_poses_all",
             [](RopeBatch& b, py::object pos, py::object quat, py::object teleport) {
                 b.set_targets(-1, pos, quat, teleport, "set_mesh_poses_all", b.num_meshes());
             },
             py::arg("pos"), py::arg("quat"), py::arg("teleport") = false)
        .def("clear_meshes", &RopeBatch::clear_meshes)
        .def("enable_contact_forces", &RopeBatch::enable_contact_forces, py::arg("on") = true,
             py::arg("deterministic") = py::none())
        .def_property_readonly("deterministic_forces", &RopeBatch::deterministic_forces)
        .def("set_gravity", &RopeBatch::set_gravity, py::arg("gravity"))
        .def("inv_masses", &RopeBatch::inv_masses)
        .def("set_masses", [](RopeBatch& b, const farray& a) { b.set_per_vertex(RS_V_MASSES, a, "set_masses", "[V] or [num_envs, V] node masses (kg)"); },
             py::arg("masses"))
        .def("set_stretch_stiffnesses",
             [](RopeBatch& b, const farray& a) {
                 b.set_per_vertex(RS_V_EA, a, "set_stretch_stiffnesses", "[V-1] or [num_envs, V-1] EA values");
             },
             py::arg("ea"))
        .def("set_bend_stiffnesses",
             [](RopeBatch& b, const farray& a) {
                 b.set_per_vertex(RS_V_EI, a, "set_bend_stiffnesses", "[V-1] or [num_envs, V-1] EI values");
             },
             py::arg("ei"))
        .def("set_dispatch", &RopeBatch::set_dispatch, py::arg("mode"), py::arg("variant") = -1,
             "0: one launch per kernel per substep; 1: one persistent frame kernel (variant -1: automatic).")
        .def_property_readonly("dispatch", &RopeBatch::dispatch)
        .def_property("rope_broad_phase", &RopeBatch::rope_broad_phase, &RopeBatch::set_rope_broad_phase,
                      "Rope-rope detection of the per-substep kernels: 'auto' (a bounding-volume hierarchy per "
                      "contact group above 128 vertices), 'scan' or 'bvh'; all give the same lists.")
        .def_property_readonly("rope_bvh_active", &RopeBatch::rope_bvh_active)
        .def("autotune", &RopeBatch::autotune, py::arg("dt"), py::arg("substeps"), py::arg("program") = true)
        .def("set_material", &RopeBatch::set_material, py::arg("ea") = py::none(), py::arg("ei") = py::none(),
             py::arg("mass") = py::none(),
             "Per-vertex EA, EI (per segment, last vertex ignored) and node masses (pinned ignored); GPU or host "
             "arrays of num_vertices float32; allowed between rollouts")
        .def("set_graph", &RopeBatch::set_graph, py::arg("on") = true,
             "Capture each frame's kernel sequence into a hipGraph and replay it while its inputs are unchanged")
        .def_property_readonly("graph_captures", &RopeBatch::graph_captures)
        .def("tape_begin", &RopeBatch::tape_begin, py::arg("max_frames"),
             "Record a checkpoint at every following step (up to max_frames) for adjoint()")
        .def("tape_end", &RopeBatch::tape_end)
        .def("tape_frames", &RopeBatch::tape_frames)
        .def("adjoint", &RopeBatch::adjoint, py::arg("dl_dx"), py::arg("g_ea") = py::none(),
             py::arg("g_ei") = py::none(), py::arg("g_mass") = py::none(), py::arg("g_p0") = py::none(),
             py::arg("g_v0") = py::none(), py::arg("g_p1") = py::none(),
             "Backward pass over the recorded frames: dl_dx [frames, num_vertices, 3] in, gradients written into "
             "the given GPU arrays; returns (dL/d drag, recompute mismatches)")
        .def("set_kernel_timing", [](RopeBatch& b, bool on) { check(rs_set_kernel_timing(b.b(), on ? 1 : 0)); },
             py::arg("on") = true)
        .def("kernel_time",
             [](RopeBatch& b) {
                 double ms = 0;
                 int n = 0;
                 check(rs_kernel_time(b.b(), &ms, &n));
                 return py::make_tuple(ms, n);
             },
             "(summed milliseconds, launches) of the frame kernels since the last call, timed with HIP events on "
             "the solver stream.")
        .def("chunk_stats",
             [](RopeBatch& b, bool on) {
                 long long h[4] = {0, 0, 0, 0};
                 check(rs_set_chunk_stats(b.b(), on ? 1 : 0, h));
                 return py::make_tuple(h[0], h[1], h[2], h[3]);
             },
             py::arg("on") = true)
        .def_property("drag", [](RopeBatch& b) { return b.param(RS_P_DRAG); }, [](RopeBatch& b, float v) { b.set_param(RS_P_DRAG, v); })
        .def_property("drag_quad", [](RopeBatch& b) { return b.param(RS_P_DRAG_QUAD); },
                      [](RopeBatch& b, float v) { b.set_param(RS_P_DRAG_QUAD, v); })
        .def_property("drag_quad_t", [](RopeBatch& b) { return b.param(RS_P_DRAG_QUAD_T); },
                      [](RopeBatch& b, float v) { b.set_param(RS_P_DRAG_QUAD_T, v); })
        .def_property("damping", [](RopeBatch& b) { return b.param(RS_P_DAMPING); },
                      [](RopeBatch& b, float v) { b.set_param(RS_P_DAMPING, v); })
        .def_property("num_itr", [](RopeBatch& b) { return (int)b.param(RS_P_NUM_ITR); },
                      [](RopeBatch& b, int v) { b.set_param(RS_P_NUM_ITR, v); })
        .def_property_readonly("time", [](RopeBatch& b) { return b.param(RS_P_TIME); })
        .def_property_readonly("last_error", &RopeBatch::last_error)
        .def_property_readonly("last_warning", &RopeBatch::last_warning)
        .def_property_readonly("num_envs", [](RopeBatch& b) { return b.N; })
        .def_property_readonly("num_vertices", [](RopeBatch& b) { return b.nv; })
        .def_property_readonly("verts_per_rope", [](RopeBatch& b) { return b.V; })
        .def_property_readonly("offsets", [](RopeBatch& b) {
            py::array_t<int> a(b.N + 1);
            for (int i = 0; i <= b.N; i++) a.mutable_data()[i] = b.off[i];
            return a;
        })
        .def_property_readonly("fixed_root_orientation", [](RopeBatch& b) { return b.fixRoot; })
        .def_property_readonly("_handle", [](RopeBatch& b) { return (uintptr_t)b.b(); });
}
