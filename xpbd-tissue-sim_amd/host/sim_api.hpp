// C++ host layer over the C-ABI (include/xpbd_hip.h) that mirrors the
// reference's simulation API for the XPBD hot path: config objects with the
// reference's option names, an XPBD mesh object whose solver runs on the GPU,
// and the simulation loop.  Reference interfaces mirrored:
//   Config::SimulationConfig        include/config/simulation/SimulationConfig.hpp:79-89
//   Config::XPBDMeshObjectConfig    include/config/simobject/XPBDMeshObjectConfig.hpp:26-71,
//                                   ObjectConfig.hpp:28-33, MeshObjectConfig.hpp:15-22,
//                                   ElasticMaterialConfig.hpp:20-24
//   Solver::XPBDGPUSolver           include/solver/xpbd_solver/XPBDSolver.hpp:47-135 (the
//                                   ParallelJacobi slot, XPBDTypedefs.hpp:96-101)
//   Sim::XPBDMeshObject             include/simobject/XPBDMeshObjectBase.hpp:96-212,
//                                   XPBDMeshObject.hpp:96-177, MeshObject.hpp:35-60
//   Sim::Simulation                 src/simulation/Simulation.cpp:197-276
//   Sim::BeamSimulation             src/simulation/BeamSimulation.cpp:18-56
// Errors from the device library become Sim::SimError exceptions carrying the
// XPBD_E_* code and message.
#pragma once
#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

#include "xpbd_hip.h"
#include "yaml_lite.hpp"

namespace Sim {

using Vec3r = std::array<double, 3>;

struct SimError : std::runtime_error {
    int code;
    SimError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

}  // namespace Sim

namespace Config {

enum class XPBDObjectSolverTypeEnum { GAUSS_SEIDEL, JACOBI, PARALLEL_JACOBI, GPU_GAUSS_SEIDEL };
enum class XPBDSolverResidualPolicyEnum { NEVER, EVERY_SUBSTEP, EVERY_ITERATION };

struct ElasticMaterialConfig {
    std::string name = "default";
    double density = 1000, E = 3e3, nu = 0.45, mu_s = 0, mu_k = 0;
};

struct RigidObjectConfig {
    std::string name, type;  // "RigidSphere" or "RigidBox"
    Sim::Vec3r position{ 0, 0, 0 }, velocity{ 0, 0, 0 }, rotation{ 0, 0, 0 };
    double radius = 0.5;
    Sim::Vec3r size{ 1, 1, 1 };
    bool fixed = true, collisions = true;
};

struct XPBDMeshObjectConfig {
    std::string name = "object", type = "XPBDMeshObject";
    std::string filename;                       // .msh (GMSH 4.1)
    std::optional<std::array<int, 3>> grid;     // or a structured grid of n cells per axis...
    double cell_size = 0.01;                    // ...with this cell edge [m]
    std::optional<double> max_size;
    std::optional<Sim::Vec3r> size;
    Sim::Vec3r position{ 0, 0, 0 }, velocity{ 0, 0, 0 }, rotation{ 0, 0, 0 };
    // elastic materials: "materials: [names]" refers to the simulation's
    // top-level material list (XPBDMeshObjectConfig.hpp:62, 112); an inline
    // "material: {...}" gives a one-material list.  Element class c uses
    // materials[c] (the last one for classes past the end, XPBDMeshObject.cpp:287-300).
    ElasticMaterialConfig material;
    std::vector<std::string> material_names;
    std::vector<ElasticMaterialConfig> materials;         // resolved list (at least one)
    std::optional<std::string> element_classes_filename;  // element-classes-filename (:63)
    bool collisions = false;
    int num_solver_iters = 1;
    int num_local_collision_iters = 0;  // num-local-collision-iters (XPBDMeshObjectConfig.hpp:68)
    XPBDObjectSolverTypeEnum solver_type = XPBDObjectSolverTypeEnum::GAUSS_SEIDEL;
    std::string constraint_type = "Stable-Neohookean-Combined";
    XPBDSolverResidualPolicyEnum residual_policy = XPBDSolverResidualPolicyEnum::NEVER;
    // vertices with the minimum coordinate along this axis are fixed (-1: none)
    int fix_min_axis = -1;
    // type "FirstOrderXPBDMeshObject" (FirstOrderXPBDMeshObjectConfig.hpp:18-59)
    bool first_order = false;
    double damping_multiplier = 1.0;        // damping-multiplier
    bool adjust_damping_to_material = false;  // adjust-damping-to-material

    static XPBDMeshObjectConfig from_node(const yaml_lite::Node& n);
};

struct SimulationConfig {
    std::string name = "simulation", description;
    double time_step = 1e-3, end_time = 10, g_accel = 9.81, collision_rate = 100, fps = 30;
    std::string sim_mode = "AFAP", visualization = "None";
    bool logging = false;
    std::string logging_output_folder = ".";
    std::vector<ElasticMaterialConfig> materials;  // top-level "materials:" list
    std::vector<XPBDMeshObjectConfig> xpbd_objects;
    std::vector<RigidObjectConfig> rigid_objects;

    static SimulationConfig from_node(const yaml_lite::Node& n);
    static SimulationConfig from_file(const std::string& path);
};

// option tables of the reference (XPBDMeshObjectConfig.hpp:26-52) plus the GPU solvers
XPBDObjectSolverTypeEnum solver_type_from_string(const std::string& s);
XPBDSolverResidualPolicyEnum residual_policy_from_string(const std::string& s);

}  // namespace Config

namespace Sim {

class XPBDMeshObject;

}  // namespace Sim

namespace Solver {

// Device solver for one mesh object: the whole second-order step (predict,
// iterations, pin reset, velocity update) runs on the GPU from a captured graph.
class XPBDGPUSolver {
public:
    XPBDGPUSolver(Sim::XPBDMeshObject* obj, int num_iter, Config::XPBDSolverResidualPolicyEnum policy,
                  Config::XPBDObjectSolverTypeEnum type);
    ~XPBDGPUSolver();
    XPBDGPUSolver(const XPBDGPUSolver&) = delete;
    XPBDGPUSolver& operator=(const XPBDGPUSolver&) = delete;

    void setup();
    void solve();                                   // one time step
    int numIterations() const { return _num_iter; }
    std::vector<double> constraintResidual() const; // r = C + alpha~ lambda, 2 per element
    double lastConstraintResidualNorm() const;      // EVERY_SUBSTEP value of the last step
    // primary residual (x - xt) / w - sum grad C^T lambda, 3 per vertex, of the
    // last step (XPBDSolver::primaryResidual, XPBDSolver.hpp:127-130), and its norm
    std::vector<double> primaryResidual() const;
    double lastPrimaryResidualNorm() const;
    double totalStrainEnergy() const;
    void setCollider(int32_t kind, const Sim::Vec3r& pos, const std::array<double, 4>& quat,
                     const Sim::Vec3r& params, const Sim::Vec3r& vel);
    void detectCollisions();
    // replaces the device attachment list: vertices[a] pulled to targets[3a..3a+2]
    void setAttachments(const std::vector<int32_t>& vertices, const std::vector<double>& targets);
    std::vector<double> positions() const;
    std::vector<double> velocities() const;
    std::vector<double> masses() const;
    xpbd_handle* handle() { return _h; }

private:
    void check(int rc) const;
    Sim::XPBDMeshObject* _obj;
    int _num_iter;
    Config::XPBDSolverResidualPolicyEnum _policy;
    Config::XPBDObjectSolverTypeEnum _type;
    xpbd_handle* _h = nullptr;
};

}  // namespace Solver

namespace Sim {

struct AABB {
    Vec3r min, max;
    Vec3r center() const { return { (min[0] + max[0]) / 2, (min[1] + max[1]) / 2, (min[2] + max[2]) / 2 }; }
};

class XPBDMeshObject {
public:
    XPBDMeshObject(const Config::XPBDMeshObjectConfig& cfg, double dt, double g);

    void setup();
    void update();          // one step of the device solver
    void velocityUpdate();  // part of the device step; kept for the call sequence
    void fixVertex(int index);
    bool vertexFixed(int index) const { return _fixed.at(index) != 0; }
    double vertexMass(int index) const { return _masses.at(index); }
    Vec3r vertexVelocity(int index) const;
    Vec3r vertex(int index) const;
    std::vector<double> lastConstraintResidual() const;
    double lastConstraintResidualNorm() const;
    std::vector<double> lastPrimaryResidual() const;
    double lastPrimaryResidualNorm() const;
    double totalStrainEnergy() const;
    AABB boundingBox() const;
    // Attachment (grasp) constraints, XPBDMeshObject.cpp:254-275: vertex v_ind is
    // pulled to *attach_pos_ptr + attachment_offset, where *attach_pos_ptr is read
    // at every update (the caller keeps it alive and moves it, e.g. a cursor).
    void addAttachmentConstraint(int v_ind, const Vec3r* attach_pos_ptr, const Vec3r& attachment_offset);
    void clearAttachmentConstraints();
    int numAttachmentConstraints() const { return (int)_att.size(); }
    // vertex of a boundary face of the tetrahedral mesh (Mesh.cpp:16-25)
    bool vertexOnSurface(int index);
    std::vector<int> getVerticesWithAxisValue(int axis, double value, double tol = 1e-9) const;
    std::vector<int> getVerticesWithY(double y) const { return getVerticesWithAxisValue(1, y); }
    int getClosestVertex(const Vec3r& p) const;
    const std::string& name() const { return _cfg.name; }
    int numVertices() const { return (int)_x0.size() / 3; }
    int numElements() const { return (int)_tets.size() / 4; }
    int numFaces() const { return (int)_faces.size() / 3; }  // surface faces (objects with collisions)
    // element class (0 without an element classes file) and vertex class, the
    // largest class of the vertex's elements (XPBDMeshObject.cpp:112-149)
    int elementClass(int e) const { return _elem_class.empty() ? 0 : _elem_class.at(e); }
    int vertexClass(int v) const { return _vert_class.empty() ? 0 : _vert_class.at(v); }
    // material of each element after the class lookup (an index into config().materials)
    int elementMaterialIndex(int e) const { return _elem_mat.at(e); }
    const Config::ElasticMaterialConfig& elementMaterial(int e) const { return _cfg.materials.at(_elem_mat.at(e)); }
    const std::vector<double>& restVertices() const { return _x0; }
    const std::vector<double>& vertices() const { return _x; }  // refreshed after each update
    Solver::XPBDGPUSolver& solver() { return *_solver; }
    const Config::XPBDMeshObjectConfig& config() const { return _cfg; }

    // filled by the solver's setup
    xpbd_mesh_desc descriptor();

private:
    void loadAndConfigureMesh();
    Config::XPBDMeshObjectConfig _cfg;
    double _dt, _g;
    std::vector<double> _x0, _x, _v0, _E, _nu, _rho, _masses;
    std::vector<int32_t> _tets, _faces, _elem_class, _vert_class, _elem_mat;
    std::vector<uint8_t> _fixed;
    struct Attachment {
        int v;
        const Vec3r* pos;
        Vec3r offset;
    };
    std::vector<Attachment> _att;
    bool _att_sent = false;  // the device holds a non-empty list
    std::vector<uint8_t> _surface;  // surface vertex flags, computed on first use
    std::unique_ptr<Solver::XPBDGPUSolver> _solver;
    friend class Solver::XPBDGPUSolver;
};

class Simulation {
public:
    explicit Simulation(const Config::SimulationConfig& cfg);
    virtual ~Simulation() = default;
    virtual void setup();
    void run(int64_t max_steps = -1);  // as fast as possible, until end-time or max_steps
    virtual void timeStep();
    double time() const { return _time; }
    int64_t steps() const { return _steps; }
    std::vector<std::unique_ptr<XPBDMeshObject>>& objects() { return _objects; }
    // per-step logging: one column per output, written after every step
    void addOutput(const std::string& name, std::function<double()> f);
    std::string logText() const { return _log; }

protected:
    // called for every object after it is loaded and before its solver is created
    virtual void configureObject(XPBDMeshObject&) {}
    Config::SimulationConfig _cfg;
    std::vector<std::unique_ptr<XPBDMeshObject>> _objects;
    double _time = 0, _last_collision_detection_time = 0, _time_between_collision_checks = 0;
    int64_t _steps = 0;
    std::vector<std::pair<std::string, std::function<double()>>> _outputs;
    std::string _log;
};

// Fixes the minimum-Y face of every object and logs tip deflection and residual.
class BeamSimulation : public Simulation {
public:
    using Simulation::Simulation;
    void setup() override;

protected:
    void configureObject(XPBDMeshObject& obj) override;
};

// Grasping without the input devices (GraspingSimulation.cpp:26-60, 150-320): a
// cursor point and a grasp radius.  toggleGrasping() attaches every non-fixed
// surface vertex within the radius of the cursor (at most 256 per object) with
// its offset from the cursor, so the attachments start satisfied, or releases
// them; moveCursor(dp) moves the point the attachments follow.  The cursor
// moves by cursorVelocity() * dt at the start of every step.
class GraspingSimulation : public Simulation {
public:
    using Simulation::Simulation;
    static constexpr int kMaxAttachPerObject = 256;
    void timeStep() override;
    void setCursor(const Vec3r& p) { _cursor = p; }
    const Vec3r& cursorPosition() const { return _cursor; }
    void setCursorVelocity(const Vec3r& v) { _cursor_vel = v; }
    const Vec3r& cursorVelocity() const { return _cursor_vel; }
    void setGraspRadius(double r) { _grasp_radius = r; }
    double graspRadius() const { return _grasp_radius; }
    void moveCursor(const Vec3r& dp);
    void toggleGrasping();
    bool grasping() const { return _grasping; }
    // (object index, vertex) of every attachment made by the last grasp
    const std::vector<std::pair<int, int>>& graspedVertices() const { return _grasped; }

private:
    Vec3r _cursor{ 0, 0, 0 }, _cursor_vel{ 0, 0, 0 };
    double _grasp_radius = 0.1;  // grasp-radius default (GraspingSimulationConfig.hpp:25)
    bool _grasping = false;
    std::vector<std::pair<int, int>> _grasped;
};

}  // namespace Sim
