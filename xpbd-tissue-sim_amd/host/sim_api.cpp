#include "sim_api.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <sstream>

namespace Config {

XPBDObjectSolverTypeEnum solver_type_from_string(const std::string& s)
{
    if (s == "Gauss-Seidel") return XPBDObjectSolverTypeEnum::GAUSS_SEIDEL;
    if (s == "Jacobi") return XPBDObjectSolverTypeEnum::JACOBI;
    if (s == "Parallel-Jacobi") return XPBDObjectSolverTypeEnum::PARALLEL_JACOBI;
    if (s == "GPU-Gauss-Seidel") return XPBDObjectSolverTypeEnum::GPU_GAUSS_SEIDEL;
    throw Sim::SimError(XPBD_E_INVALID_ARG, "unknown solver-type '" + s + "'");
}

XPBDSolverResidualPolicyEnum residual_policy_from_string(const std::string& s)
{
    if (s == "Never") return XPBDSolverResidualPolicyEnum::NEVER;
    if (s == "Every-Substep") return XPBDSolverResidualPolicyEnum::EVERY_SUBSTEP;
    if (s == "Every-Iteration") return XPBDSolverResidualPolicyEnum::EVERY_ITERATION;
    throw Sim::SimError(XPBD_E_INVALID_ARG, "unknown residual-policy '" + s + "'");
}

namespace {
Sim::Vec3r vec3(const yaml_lite::Node& n, const char* what)
{
    const std::vector<double> v = n.as_doubles();
    if (v.size() != 3) throw Sim::SimError(XPBD_E_INVALID_ARG, std::string(what) + " needs 3 values");
    return { v[0], v[1], v[2] };
}
template <typename T, typename F>
void opt(const yaml_lite::Node& n, const char* key, T& out, F conv)
{
    if (const yaml_lite::Node* c = n.find(key))
        if (c->kind != yaml_lite::Node::Null) out = conv(*c);
}
}  // namespace

// ElasticMaterialConfig.hpp:20-24 (name, density, E, nu, mu-s, mu-k)
static ElasticMaterialConfig material_from_node(const yaml_lite::Node& m)
{
    ElasticMaterialConfig c;
    auto S = [](const yaml_lite::Node& x) { return x.as_string(); };
    auto D = [](const yaml_lite::Node& x) { return x.as_double(); };
    opt(m, "name", c.name, S);
    opt(m, "density", c.density, D);
    opt(m, "E", c.E, D);
    opt(m, "nu", c.nu, D);
    opt(m, "mu-s", c.mu_s, D);
    opt(m, "mu-k", c.mu_k, D);
    return c;
}

XPBDMeshObjectConfig XPBDMeshObjectConfig::from_node(const yaml_lite::Node& n)
{
    XPBDMeshObjectConfig c;
    auto S = [](const yaml_lite::Node& x) { return x.as_string(); };
    auto D = [](const yaml_lite::Node& x) { return x.as_double(); };
    auto B = [](const yaml_lite::Node& x) { return x.as_bool(); };
    opt(n, "name", c.name, S);
    opt(n, "type", c.type, S);
    if (c.type == "FirstOrderXPBDMeshObject") {
        c.first_order = true;
        opt(n, "damping-multiplier", c.damping_multiplier, D);
        opt(n, "adjust-damping-to-material", c.adjust_damping_to_material, B);
        if (!(c.damping_multiplier > 0))
            throw Sim::SimError(XPBD_E_INVALID_ARG, "object '" + c.name + "': damping-multiplier must be > 0");
    } else if (c.type != "XPBDMeshObject") {
        throw Sim::SimError(XPBD_E_UNSUPPORTED, "object '" + c.name + "': type " + c.type +
                                                    " is not on the device path (XPBDMeshObject, "
                                                    "FirstOrderXPBDMeshObject)");
    }
    opt(n, "filename", c.filename, S);
    if (const yaml_lite::Node* g = n.find("grid")) {
        const std::vector<double> v = g->as_doubles();
        if (v.size() != 3) throw Sim::SimError(XPBD_E_INVALID_ARG, "grid needs 3 cell counts");
        c.grid = std::array<int, 3>{ (int)v[0], (int)v[1], (int)v[2] };
    }
    opt(n, "cell-size", c.cell_size, D);
    if (const yaml_lite::Node* m = n.find("max-size")) c.max_size = m->as_double();
    if (const yaml_lite::Node* m = n.find("size")) c.size = vec3(*m, "size");
    if (const yaml_lite::Node* m = n.find("position")) c.position = vec3(*m, "position");
    if (const yaml_lite::Node* m = n.find("velocity")) c.velocity = vec3(*m, "velocity");
    if (const yaml_lite::Node* m = n.find("rotation")) c.rotation = vec3(*m, "rotation");
    if (const yaml_lite::Node* m = n.find("material")) c.material = material_from_node(*m);
    if (const yaml_lite::Node* m = n.find("materials")) {
        if (m->kind != yaml_lite::Node::Seq) throw Sim::SimError(XPBD_E_INVALID_ARG, "materials must be a list of names");
        for (const yaml_lite::Node& x : m->seq) c.material_names.push_back(x.as_string());
    }
    opt(n, "element-classes-filename", c.element_classes_filename, S);
    opt(n, "collisions", c.collisions, B);
    if (const yaml_lite::Node* m = n.find("num-solver-iters")) c.num_solver_iters = (int)m->as_int();
    if (const yaml_lite::Node* m = n.find("num-local-collision-iters"))
        c.num_local_collision_iters = (int)m->as_int();
    if (const yaml_lite::Node* m = n.find("solver-type")) c.solver_type = solver_type_from_string(m->as_string());
    opt(n, "constraint-type", c.constraint_type, S);
    if (c.constraint_type != "Stable-Neohookean-Combined")
        throw Sim::SimError(XPBD_E_UNSUPPORTED, "constraint-type must be Stable-Neohookean-Combined on the device");
    if (const yaml_lite::Node* m = n.find("residual-policy"))
        c.residual_policy = residual_policy_from_string(m->as_string());
    if (const yaml_lite::Node* m = n.find("fix-min-axis")) {
        const std::string a = m->as_string();
        c.fix_min_axis = a == "x" ? 0 : a == "y" ? 1 : a == "z" ? 2 : a == "none" ? -1 : -2;
        if (c.fix_min_axis == -2) throw Sim::SimError(XPBD_E_INVALID_ARG, "fix-min-axis must be x, y, z or none");
    }
    if (c.filename.empty() && !c.grid)
        throw Sim::SimError(XPBD_E_INVALID_ARG, "object '" + c.name + "' needs a filename or a grid");
    if (c.num_solver_iters < 0) throw Sim::SimError(XPBD_E_INVALID_ARG, "num-solver-iters < 0");
    if (c.num_local_collision_iters < 0)
        throw Sim::SimError(XPBD_E_INVALID_ARG, "num-local-collision-iters < 0");
    return c;
}

SimulationConfig SimulationConfig::from_node(const yaml_lite::Node& n)
{
    SimulationConfig c;
    auto S = [](const yaml_lite::Node& x) { return x.as_string(); };
    auto D = [](const yaml_lite::Node& x) { return x.as_double(); };
    auto B = [](const yaml_lite::Node& x) { return x.as_bool(); };
    opt(n, "name", c.name, S);
    opt(n, "description", c.description, S);
    opt(n, "time-step", c.time_step, D);
    opt(n, "end-time", c.end_time, D);
    opt(n, "g-accel", c.g_accel, D);
    opt(n, "collision-rate", c.collision_rate, D);
    opt(n, "fps", c.fps, D);
    opt(n, "sim-mode", c.sim_mode, S);
    opt(n, "visualization", c.visualization, S);
    opt(n, "logging", c.logging, B);
    opt(n, "logging-output-folder", c.logging_output_folder, S);
    if (!(c.time_step > 0)) throw Sim::SimError(XPBD_E_INVALID_ARG, "time-step must be > 0");
    if (const yaml_lite::Node* mats = n.find("materials")) {
        if (mats->kind != yaml_lite::Node::Seq) throw Sim::SimError(XPBD_E_INVALID_ARG, "materials must be a list");
        for (const yaml_lite::Node& m : mats->seq) c.materials.push_back(material_from_node(m));
    }
    if (const yaml_lite::Node* objs = n.find("objects")) {
        if (objs->kind != yaml_lite::Node::Seq) throw Sim::SimError(XPBD_E_INVALID_ARG, "objects must be a list");
        for (const yaml_lite::Node& o : objs->seq) {
            const std::string type = o.has("type") ? o["type"].as_string() : "";
            if (type == "RigidSphere" || type == "RigidBox") {
                RigidObjectConfig r;
                r.type = type;
                opt(o, "name", r.name, S);
                if (const yaml_lite::Node* m = o.find("position")) r.position = vec3(*m, "position");
                if (const yaml_lite::Node* m = o.find("velocity")) r.velocity = vec3(*m, "velocity");
                if (const yaml_lite::Node* m = o.find("rotation")) r.rotation = vec3(*m, "rotation");
                if (const yaml_lite::Node* m = o.find("size")) r.size = vec3(*m, "size");
                opt(o, "radius", r.radius, D);
                opt(o, "fixed", r.fixed, B);
                opt(o, "collisions", r.collisions, B);
                c.rigid_objects.push_back(r);
            } else {
                c.xpbd_objects.push_back(XPBDMeshObjectConfig::from_node(o));
            }
        }
    }
    // resolve each object's material names against the top-level list
    for (XPBDMeshObjectConfig& o : c.xpbd_objects) {
        o.materials.clear();
        for (const std::string& name : o.material_names) {
            auto it = std::find_if(c.materials.begin(), c.materials.end(),
                                   [&](const ElasticMaterialConfig& m) { return m.name == name; });
            if (it == c.materials.end())
                throw Sim::SimError(XPBD_E_INVALID_ARG, "object '" + o.name + "': no material named '" + name + "'");
            o.materials.push_back(*it);
        }
        if (o.materials.empty()) o.materials.push_back(o.material);
    }
    return c;
}

SimulationConfig SimulationConfig::from_file(const std::string& path)
{
    try {
        return from_node(yaml_lite::parse_file(path));
    } catch (const yaml_lite::ParseError& e) {
        throw Sim::SimError(XPBD_E_INVALID_ARG, path + ": " + e.what());
    }
}

}  // namespace Config

// ---------------------------------------------------------------------------
namespace Solver {

using Config::XPBDObjectSolverTypeEnum;
using Config::XPBDSolverResidualPolicyEnum;

XPBDGPUSolver::XPBDGPUSolver(Sim::XPBDMeshObject* obj, int num_iter, XPBDSolverResidualPolicyEnum policy,
                             XPBDObjectSolverTypeEnum type)
    : _obj(obj), _num_iter(num_iter), _policy(policy), _type(type)
{
}

XPBDGPUSolver::~XPBDGPUSolver()
{
    if (_h) xpbd_destroy(_h);
}

void XPBDGPUSolver::check(int rc) const
{
    if (rc != XPBD_OK) throw Sim::SimError(rc, std::string("device solver: ") + xpbd_last_error(_h));
}

void XPBDGPUSolver::setup()
{
    xpbd_mesh_desc d = _obj->descriptor();
    xpbd_params p{};
    p.dt = _obj->_dt;
    p.g = _obj->_g;
    p.iters = _num_iter;
    // Gauss-Seidel objects run coloured Gauss-Seidel on the device; Jacobi and the
    // reference's Parallel-Jacobi slot run the plain-sum Jacobi
    p.solver = (_type == XPBDObjectSolverTypeEnum::GAUSS_SEIDEL || _type == XPBDObjectSolverTypeEnum::GPU_GAUSS_SEIDEL)
                   ? XPBD_SOLVER_GS_COLORED
                   : XPBD_SOLVER_JACOBI;
    p.deterministic = 1;
    p.collision_every = 0;  // detection is driven by Simulation::timeStep
    p.variant = XPBD_VARIANT_AUTO;
    p.first_order = _obj->config().first_order ? 1 : 0;
    p.damping_multiplier = _obj->config().damping_multiplier;
    p.adjust_damping = _obj->config().adjust_damping_to_material ? 1 : 0;
    p.local_collision_iters = _obj->config().collisions ? _obj->config().num_local_collision_iters : 0;
    const int rc = xpbd_create(&d, &p, &_h);
    if (rc != XPBD_OK) throw Sim::SimError(rc, std::string("xpbd_create: ") + xpbd_last_error(nullptr));
    if (_policy != XPBDSolverResidualPolicyEnum::NEVER) {
        // EVERY_SUBSTEP records both residuals (XPBDSolver.hpp:91-96)
        check(xpbd_set_residual_policy(_h, 1));
        check(xpbd_set_primary_residual(_h, 1));
    }
}

void XPBDGPUSolver::solve() { check(xpbd_step(_h, 1)); }

std::vector<double> XPBDGPUSolver::constraintResidual() const
{
    std::vector<double> r(2 * (size_t)_obj->numElements());
    double norm = 0;
    check(xpbd_constraint_residual(_h, &norm, r.data()));
    return r;
}

double XPBDGPUSolver::lastConstraintResidualNorm() const
{
    double r = 0;
    check(xpbd_last_step_residual(_h, &r));
    return r;
}

std::vector<double> XPBDGPUSolver::primaryResidual() const
{
    std::vector<double> r(3 * (size_t)_obj->numVertices());
    double norm = 0;
    check(xpbd_last_step_primary_residual(_h, &norm, r.data()));
    return r;
}

double XPBDGPUSolver::lastPrimaryResidualNorm() const
{
    double r = 0;
    check(xpbd_last_step_primary_residual(_h, &r, nullptr));
    return r;
}

double XPBDGPUSolver::totalStrainEnergy() const
{
    double e = 0;
    check(xpbd_strain_energy(_h, &e));
    return e;
}

void XPBDGPUSolver::setCollider(int32_t kind, const Sim::Vec3r& pos, const std::array<double, 4>& quat,
                                const Sim::Vec3r& params, const Sim::Vec3r& vel)
{
    const double pose[7] = { pos[0], pos[1], pos[2], quat[0], quat[1], quat[2], quat[3] };
    check(xpbd_set_sdf(_h, kind, pose, params.data()));
    check(xpbd_set_sdf_velocity(_h, vel.data()));
}

void XPBDGPUSolver::detectCollisions() { check(xpbd_detect_contacts(_h)); }

void XPBDGPUSolver::setAttachments(const std::vector<int32_t>& vertices, const std::vector<double>& targets)
{
    if (targets.size() != 3 * vertices.size()) throw Sim::SimError(XPBD_E_INVALID_ARG, "attachments: 3 target values per vertex");
    check(xpbd_set_attachments(_h, (int32_t)vertices.size(), vertices.data(), targets.data()));
}

std::vector<double> XPBDGPUSolver::positions() const
{
    std::vector<double> x(3 * (size_t)_obj->numVertices());
    check(xpbd_get_positions(_h, x.data()));
    return x;
}

std::vector<double> XPBDGPUSolver::velocities() const
{
    std::vector<double> v(3 * (size_t)_obj->numVertices());
    check(xpbd_get_velocities(_h, v.data()));
    return v;
}

std::vector<double> XPBDGPUSolver::masses() const
{
    std::vector<double> m(_obj->numVertices());
    check(xpbd_get_masses(_h, m.data()));
    return m;
}

}  // namespace Solver

// ---------------------------------------------------------------------------
namespace Sim {

XPBDMeshObject::XPBDMeshObject(const Config::XPBDMeshObjectConfig& cfg, double dt, double g) : _cfg(cfg), _dt(dt), _g(g)
{
    loadAndConfigureMesh();
}

// MeshObject.hpp:35-60: load, scale (max-size, then size), move the centre of
// mass to the position, rotate about it (XYZ Euler angles, Mesh.cpp:285-312).
void XPBDMeshObject::loadAndConfigureMesh()
{
    if (_cfg.grid) {
        const auto& g = *_cfg.grid;
        const size_t nv = (size_t)(g[0] + 1) * (g[1] + 1) * (g[2] + 1), nt = 6 * (size_t)g[0] * g[1] * g[2];
        _x0.resize(3 * nv);
        _tets.resize(4 * nt);
        const double o[3] = { 0, 0, 0 };
        const int rc = xpbd_mesh_grid(g[0], g[1], g[2], _cfg.cell_size, _cfg.cell_size, _cfg.cell_size, o, _x0.data(),
                                      _tets.data());
        if (rc != XPBD_OK) throw SimError(rc, "grid generation failed");
    } else {
        int32_t nv = 0, nt = 0, nf = 0;
        int rc = xpbd_read_msh(_cfg.filename.c_str(), &nv, &nt, &nf, nullptr, nullptr, nullptr);
        if (rc != XPBD_OK) throw SimError(rc, "cannot read mesh " + _cfg.filename);
        _x0.resize(3 * (size_t)nv);
        _tets.resize(4 * (size_t)nt);
        std::vector<int32_t> tris(3 * (size_t)std::max(nf, 1));
        rc = xpbd_read_msh(_cfg.filename.c_str(), &nv, &nt, &nf, _x0.data(), _tets.data(), tris.data());
        if (rc != XPBD_OK) throw SimError(rc, "cannot read mesh " + _cfg.filename);
    }
    const int nv = numVertices(), nt = numElements();
    auto bbox = [&]() {
        AABB b{ { 1e300, 1e300, 1e300 }, { -1e300, -1e300, -1e300 } };
        for (int i = 0; i < nv; i++)
            for (int k = 0; k < 3; k++) {
                b.min[k] = std::min(b.min[k], _x0[3 * i + k]);
                b.max[k] = std::max(b.max[k], _x0[3 * i + k]);
            }
        return b;
    };
    if (_cfg.max_size) {
        const AABB b = bbox();
        const double ext = std::max({ b.max[0] - b.min[0], b.max[1] - b.min[1], b.max[2] - b.min[2] });
        const double s = *_cfg.max_size / ext;
        for (double& v : _x0) v *= s;
    }
    if (_cfg.size) {
        const AABB b = bbox();
        for (int k = 0; k < 3; k++) {
            const double ext = b.max[k] - b.min[k];
            const double s = ext != 0 ? (*_cfg.size)[k] / ext : 1.0;
            for (int i = 0; i < nv; i++) _x0[3 * i + k] *= s;
        }
    }
    // volume centroid (the reference integrates over the surface; equal for a closed mesh)
    double vol = 0, c[3] = { 0, 0, 0 };
    for (int e = 0; e < nt; e++) {
        const int32_t* t = &_tets[4 * (size_t)e];
        double a[3], b[3], d[3];
        for (int k = 0; k < 3; k++) {
            a[k] = _x0[3 * t[1] + k] - _x0[3 * t[0] + k];
            b[k] = _x0[3 * t[2] + k] - _x0[3 * t[0] + k];
            d[k] = _x0[3 * t[3] + k] - _x0[3 * t[0] + k];
        }
        const double v = std::fabs(a[0] * (b[1] * d[2] - b[2] * d[1]) - a[1] * (b[0] * d[2] - b[2] * d[0]) +
                                   a[2] * (b[0] * d[1] - b[1] * d[0])) / 6.0;
        vol += v;
        for (int k = 0; k < 3; k++)
            c[k] += v * 0.25 * (_x0[3 * t[0] + k] + _x0[3 * t[1] + k] + _x0[3 * t[2] + k] + _x0[3 * t[3] + k]);
    }
    for (int k = 0; k < 3; k++) c[k] /= vol;
    for (int i = 0; i < nv; i++)
        for (int k = 0; k < 3; k++) _x0[3 * i + k] += _cfg.position[k] - c[k];
    if (_cfg.rotation[0] != 0 || _cfg.rotation[1] != 0 || _cfg.rotation[2] != 0) {
        const double x = _cfg.rotation[0] * M_PI / 180, y = _cfg.rotation[1] * M_PI / 180, z = _cfg.rotation[2] * M_PI / 180;
        const double R[3][3] = {
            { std::cos(y) * std::cos(z), std::sin(x) * std::sin(y) * std::cos(z) - std::cos(x) * std::sin(z),
              std::cos(x) * std::sin(y) * std::cos(z) + std::sin(x) * std::sin(z) },
            { std::cos(y) * std::sin(z), std::sin(x) * std::sin(y) * std::sin(z) + std::cos(x) * std::cos(z),
              std::cos(x) * std::sin(y) * std::sin(z) - std::sin(x) * std::cos(z) },
            { -std::sin(y), std::sin(x) * std::cos(y), std::cos(x) * std::cos(y) } };
        for (int i = 0; i < nv; i++) {
            double p[3];
            for (int k = 0; k < 3; k++) p[k] = _x0[3 * i + k] - _cfg.position[k];
            for (int r = 0; r < 3; r++)
                _x0[3 * i + r] = _cfg.position[r] + R[r][0] * p[0] + R[r][1] * p[1] + R[r][2] * p[2];
        }
    }
    _x = _x0;
    _fixed.assign(nv, 0);
    _v0.assign(3 * (size_t)nv, 0.0);
    for (int i = 0; i < nv; i++)
        for (int k = 0; k < 3; k++) _v0[3 * i + k] = _cfg.velocity[k];
    // element classes and materials, XPBDMeshObject.cpp:107-149 and 287-300
    if (_cfg.materials.empty()) _cfg.materials.push_back(_cfg.material);
    _elem_class.clear();
    _vert_class.clear();
    if (_cfg.element_classes_filename) {
        const char* path = _cfg.element_classes_filename->c_str();
        int32_t n = 0;
        int rc = xpbd_read_element_classes(path, 0, nullptr, &n);
        if (rc != XPBD_OK) throw SimError(rc, "cannot read element classes " + *_cfg.element_classes_filename);
        if (n != nt)
            throw SimError(XPBD_E_INVALID_ARG, "element classes file " + *_cfg.element_classes_filename + " has " +
                                                   std::to_string(n) + " entries for " + std::to_string(nt) + " elements");
        _elem_class.resize(nt);
        rc = xpbd_read_element_classes(path, nt, _elem_class.data(), &n);
        if (rc != XPBD_OK) throw SimError(rc, "cannot read element classes " + *_cfg.element_classes_filename);
        _vert_class.assign(nv, 0);
        for (int e = 0; e < nt; e++)
            for (int j = 0; j < 4; j++) {
                int& vc = _vert_class[_tets[4 * (size_t)e + j]];
                vc = std::max(vc, _elem_class[e]);
            }
    }
    const int nmat = (int)_cfg.materials.size();
    _elem_mat.assign(nt, 0);
    _E.resize(nt);
    _nu.resize(nt);
    _rho.resize(nt);
    bool warned = false;
    for (int e = 0; e < nt; e++) {
        int m = elementClass(e);
        if (static_cast<unsigned>(m) >= static_cast<unsigned>(nmat)) {
            if (!warned)
                std::fprintf(stderr, "WARNING: only %d materials were specified, but element %d has class %d; "
                                     "such elements use the last material\n", nmat, e, m);
            warned = true;
            m = nmat - 1;
        }
        _elem_mat[e] = m;
        _E[e] = _cfg.materials[m].E;
        _nu[e] = _cfg.materials[m].nu;
        _rho[e] = _cfg.materials[m].density;
    }
    if (_cfg.fix_min_axis >= 0) {
        const AABB b = bbox();
        for (int v : getVerticesWithAxisValue(_cfg.fix_min_axis, b.min[_cfg.fix_min_axis])) _fixed[v] = 1;
    }
}

xpbd_mesh_desc XPBDMeshObject::descriptor()
{
    if (_cfg.collisions && _faces.empty()) {
        int32_t nf = 0;
        std::vector<int32_t> f(12 * (size_t)numElements());
        const int rc = xpbd_mesh_surface(numVertices(), numElements(), _tets.data(), &nf, f.data());
        if (rc != XPBD_OK) throw SimError(rc, "surface extraction failed");
        f.resize(3 * (size_t)nf);
        _faces = f;
    }
    xpbd_mesh_desc d{};
    d.num_vertices = numVertices();
    d.num_tets = numElements();
    d.x0 = _x0.data();
    d.v0 = _v0.data();
    d.tets = _tets.data();
    d.E = _E.data();
    d.nu = _nu.data();
    d.rho = _rho.data();
    d.pinned = _fixed.data();
    d.num_faces = (int32_t)(_faces.size() / 3);
    d.faces = _faces.empty() ? nullptr : _faces.data();
    return d;
}

void XPBDMeshObject::setup()
{
    _solver = std::make_unique<Solver::XPBDGPUSolver>(this, _cfg.num_solver_iters, _cfg.residual_policy,
                                                      _cfg.solver_type);
    _solver->setup();
    _masses = _solver->masses();
}

void XPBDMeshObject::update()
{
    if (!_att.empty() || _att_sent) {
        // targets follow the attached points, read once per step like the
        // reference's projector reads *attach_pos_ptr
        std::vector<int32_t> v(_att.size());
        std::vector<double> t(3 * _att.size());
        for (size_t a = 0; a < _att.size(); a++) {
            v[a] = _att[a].v;
            for (int k = 0; k < 3; k++) t[3 * a + k] = (*_att[a].pos)[k] + _att[a].offset[k];
        }
        _solver->setAttachments(v, t);
        _att_sent = !_att.empty();
    }
    _solver->solve();
    _x = _solver->positions();
}

void XPBDMeshObject::velocityUpdate() {}

void XPBDMeshObject::addAttachmentConstraint(int v_ind, const Vec3r* attach_pos_ptr, const Vec3r& attachment_offset)
{
    if (v_ind < 0 || v_ind >= numVertices()) throw SimError(XPBD_E_INVALID_ARG, "attachment vertex out of range");
    if (!attach_pos_ptr) throw SimError(XPBD_E_INVALID_ARG, "attachment position pointer is null");
    for (const Attachment& a : _att)
        if (a.v == v_ind) throw SimError(XPBD_E_INVALID_ARG, "a vertex may carry one attachment");
    _att.push_back({ v_ind, attach_pos_ptr, attachment_offset });
}

void XPBDMeshObject::clearAttachmentConstraints() { _att.clear(); }

bool XPBDMeshObject::vertexOnSurface(int index)
{
    if (_surface.empty()) {
        std::vector<int32_t> f = _faces;
        if (f.empty()) {
            int32_t nf = 0;
            f.resize(12 * (size_t)numElements());
            const int rc = xpbd_mesh_surface(numVertices(), numElements(), _tets.data(), &nf, f.data());
            if (rc != XPBD_OK) throw SimError(rc, "surface extraction failed");
            f.resize(3 * (size_t)nf);
        }
        _surface.assign(numVertices(), 0);
        for (int32_t v : f) _surface[v] = 1;
    }
    return _surface.at(index) != 0;
}

void XPBDMeshObject::fixVertex(int index)
{
    if (_solver) throw SimError(XPBD_E_STATE, "fixVertex must be called before setup()");
    _fixed.at(index) = 1;
}

Vec3r XPBDMeshObject::vertexVelocity(int index) const
{
    const std::vector<double> v = _solver->velocities();
    return { v[3 * index], v[3 * index + 1], v[3 * index + 2] };
}

Vec3r XPBDMeshObject::vertex(int index) const { return { _x[3 * index], _x[3 * index + 1], _x[3 * index + 2] }; }

std::vector<double> XPBDMeshObject::lastConstraintResidual() const { return _solver->constraintResidual(); }

double XPBDMeshObject::lastConstraintResidualNorm() const { return _solver->lastConstraintResidualNorm(); }
std::vector<double> XPBDMeshObject::lastPrimaryResidual() const { return _solver->primaryResidual(); }
double XPBDMeshObject::lastPrimaryResidualNorm() const { return _solver->lastPrimaryResidualNorm(); }

double XPBDMeshObject::totalStrainEnergy() const { return _solver->totalStrainEnergy(); }

AABB XPBDMeshObject::boundingBox() const
{
    AABB b{ { 1e300, 1e300, 1e300 }, { -1e300, -1e300, -1e300 } };
    for (int i = 0; i < numVertices(); i++)
        for (int k = 0; k < 3; k++) {
            b.min[k] = std::min(b.min[k], _x[3 * i + k]);
            b.max[k] = std::max(b.max[k], _x[3 * i + k]);
        }
    return b;
}

std::vector<int> XPBDMeshObject::getVerticesWithAxisValue(int axis, double value, double tol) const
{
    std::vector<int> out;
    for (int i = 0; i < numVertices(); i++)
        if (std::fabs(_x[3 * i + axis] - value) <= tol) out.push_back(i);
    return out;
}

int XPBDMeshObject::getClosestVertex(const Vec3r& p) const
{
    int best = -1;
    double bd = std::numeric_limits<double>::infinity();
    for (int i = 0; i < numVertices(); i++) {
        double d = 0;
        for (int k = 0; k < 3; k++) d += (_x[3 * i + k] - p[k]) * (_x[3 * i + k] - p[k]);
        if (d < bd) {
            bd = d;
            best = i;
        }
    }
    return best;
}

Simulation::Simulation(const Config::SimulationConfig& cfg) : _cfg(cfg)
{
    _time_between_collision_checks = 1.0 / cfg.collision_rate;
}

namespace {
// XYZ Euler angles in degrees (Mesh.cpp:285-304 convention, R = Rz Ry Rx) -> quaternion (x, y, z, w)
std::array<double, 4> euler_xyz_quat(const Vec3r& deg)
{
    const double hx = deg[0] * M_PI / 360, hy = deg[1] * M_PI / 360, hz = deg[2] * M_PI / 360;
    const double cx = std::cos(hx), sx = std::sin(hx), cy = std::cos(hy), sy = std::sin(hy), cz = std::cos(hz),
                 sz = std::sin(hz);
    return { sx * cy * cz - cx * sy * sz, cx * sy * cz + sx * cy * sz, cx * cy * sz - sx * sy * cz,
             cx * cy * cz + sx * sy * sz };
}
}  // namespace

void Simulation::setup()
{
    for (const auto& oc : _cfg.xpbd_objects) {
        auto obj = std::make_unique<XPBDMeshObject>(oc, _cfg.time_step, _cfg.g_accel);
        configureObject(*obj);
        obj->setup();
        _objects.push_back(std::move(obj));
    }
    // colliders: the device holds one kinematic SDF per object (a rigid sphere
    // or box moving at its configured velocity)
    std::vector<const Config::RigidObjectConfig*> rig;
    for (const auto& r : _cfg.rigid_objects)
        if (r.collisions) rig.push_back(&r);
    for (auto& o : _objects) {
        if (!o->config().collisions || rig.empty()) continue;
        if (rig.size() > 1)
            throw SimError(XPBD_E_UNSUPPORTED, "object '" + o->name() + "': one colliding rigid object per deformable "
                                                                        "object is supported on the device");
        const Config::RigidObjectConfig& r = *rig[0];
        const bool sphere = r.type == "RigidSphere";
        const Vec3r prm = sphere ? Vec3r{ r.radius, 0, 0 } : r.size;
        o->solver().setCollider(sphere ? XPBD_SDF_SPHERE : XPBD_SDF_BOX, r.position, euler_xyz_quat(r.rotation), prm,
                                r.velocity);
    }
}

void Simulation::addOutput(const std::string& name, std::function<double()> f) { _outputs.emplace_back(name, f); }

void Simulation::run(int64_t max_steps)
{
    if (!_outputs.empty()) {
        _log += "time";
        for (const auto& o : _outputs) _log += " " + o.first;
        _log += "\n";
    }
    while (_time < _cfg.end_time && (max_steps < 0 || _steps < max_steps)) timeStep();
}

// Simulation.cpp:197-276: collision detection when more than
// 1/collision-rate has passed since the last one (strict '>'), then every
// object's update and velocity update, logging, and t += dt.
void Simulation::timeStep()
{
    const bool detect = _time - _last_collision_detection_time > _time_between_collision_checks;
    if (detect)
        for (auto& o : _objects)
            if (o->config().collisions) o->solver().detectCollisions();
    for (auto& o : _objects) o->update();
    for (auto& o : _objects) o->velocityUpdate();
    if (detect) _last_collision_detection_time = _time;
    if (!_outputs.empty()) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.9g", _time);
        _log += buf;
        for (const auto& o : _outputs) {
            std::snprintf(buf, sizeof buf, " %.9g", o.second());
            _log += buf;
        }
        _log += "\n";
    }
    _time += _cfg.time_step;
    _steps++;
}

// fix the minimum-Y side of each object before its solver is created (BeamSimulation.cpp:25-33)
void BeamSimulation::configureObject(XPBDMeshObject& obj)
{
    const AABB b = obj.boundingBox();
    for (int v : obj.getVerticesWithY(b.min[1])) obj.fixVertex(v);
}

void BeamSimulation::setup()
{
    Simulation::setup();
    for (auto& o : _objects) {
        XPBDMeshObject* p = o.get();
        const AABB b = p->boundingBox();
        const Vec3r c = b.center();
        const int tip = p->getClosestVertex({ c[0], b.max[1], c[2] });
        const double z0 = p->vertex(tip)[2];
        addOutput(p->name() + "_Z-deflection[m]", [p, tip, z0]() { return z0 - p->vertex(tip)[2]; });
        if (p->config().residual_policy != Config::XPBDSolverResidualPolicyEnum::NEVER) {
            addOutput(p->name() + "_Constraint-Residual", [p]() { return p->lastConstraintResidualNorm(); });
            addOutput(p->name() + "_Primary-Residual", [p]() { return p->lastPrimaryResidualNorm(); });
        }
    }
}

}  // namespace Sim

namespace Sim {

void GraspingSimulation::moveCursor(const Vec3r& dp)
{
    for (int k = 0; k < 3; k++) _cursor[k] += dp[k];
}

void GraspingSimulation::timeStep()
{
    const double dt = _cfg.time_step;
    moveCursor({ _cursor_vel[0] * dt, _cursor_vel[1] * dt, _cursor_vel[2] * dt });
    Simulation::timeStep();
}

void GraspingSimulation::toggleGrasping()
{
    // GraspingSimulation.cpp:177-318: a clean slate, then every non-fixed surface
    // vertex within the radius, vertex order, capped per object
    for (auto& o : _objects) o->clearAttachmentConstraints();
    _grasped.clear();
    if (!_grasping) {
        for (size_t oi = 0; oi < _objects.size(); oi++) {
            XPBDMeshObject& o = *_objects[oi];
            int attached = 0;
            for (int v = 0; v < o.numVertices(); v++) {
                if (o.vertexFixed(v) || !o.vertexOnSurface(v)) continue;
                const Vec3r x = o.vertex(v);
                const double d = std::sqrt((x[0] - _cursor[0]) * (x[0] - _cursor[0]) +
                                           (x[1] - _cursor[1]) * (x[1] - _cursor[1]) +
                                           (x[2] - _cursor[2]) * (x[2] - _cursor[2]));
                if (d <= _grasp_radius && attached < kMaxAttachPerObject) {
                    o.addAttachmentConstraint(v, &_cursor, { x[0] - _cursor[0], x[1] - _cursor[1], x[2] - _cursor[2] });
                    _grasped.emplace_back((int)oi, v);
                    attached++;
                }
            }
        }
    }
    _grasping = !_grasping;
}

}  // namespace Sim
