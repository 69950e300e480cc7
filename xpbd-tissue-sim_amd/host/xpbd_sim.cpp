// Command-line driver of the host layer: runs a simulation config on the GPU
// as fast as possible and writes the per-step log.
//
//   xpbd_sim CONFIG.yaml [--beam] [--steps N] [--dump-positions FILE]
//            [--dump-order FILE] [--dump-face-rank FILE] [--dump-rest FILE] [--print-config]
//            [--grasp R [--cursor X,Y,Z] [--cursor-velocity VX,VY,VZ] [--dump-attachments FILE]]
//
// --beam            BeamSimulation set-up (fix the minimum-Y side, log tip
//                   deflection and constraint residual)
// --steps N         stop after N steps (default: end-time / time-step)
// --dump-positions  final vertex positions of object 0, raw float64 (V x 3)
// --dump-order      the device's element order of object 0, raw int32 (T)
// --dump-face-rank  the device's contact visiting rank of object 0's faces, raw int32
// --dump-rest       rest positions of object 0 after placement, raw float64 (V x 3)
// --print-config    parse the config, print a summary, touch no GPU
// --load-meshes     with --print-config: also load and place each mesh (host only)
// --solver-type S   override every object's solver-type (a config solver-type name)
// --autotune        time the kernel variants of every object after set-up and keep
//                   the fastest (xpbd_autotune)
// --tune-cache FILE with --autotune: reuse (and append to) the tuning cache FILE,
//                   so a run takes the kernel shape an earlier run chose for the
//                   same problem shape
// --eager           launch every step without the HIP graph (kernel timing mode);
//                   needed by profilers that cannot follow graph launches
// --copy-calibration MB
//                   first copy MB mebibytes device-to-device once: a transfer of
//                   known size to scale hardware byte counters against
// --grasp R         GraspingSimulation: after set-up, grasp every non-fixed surface
//                   vertex within R of the cursor
// --cursor X,Y,Z    cursor start (default: centre of object 0's top bounding-box face)
// --cursor-velocity VX,VY,VZ
//                   cursor motion per second of simulated time while grasping
// --dump-attachments
//                   object 0's grasped vertex ids, raw int32
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "sim_api.hpp"

static const char* solver_name(Config::XPBDObjectSolverTypeEnum t)
{
    switch (t) {
    case Config::XPBDObjectSolverTypeEnum::GAUSS_SEIDEL: return "Gauss-Seidel";
    case Config::XPBDObjectSolverTypeEnum::JACOBI: return "Jacobi";
    case Config::XPBDObjectSolverTypeEnum::PARALLEL_JACOBI: return "Parallel-Jacobi";
    case Config::XPBDObjectSolverTypeEnum::GPU_GAUSS_SEIDEL: return "GPU-Gauss-Seidel";
    }
    return "?";
}

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s CONFIG.yaml [--beam] [--steps N] [--dump-positions FILE] "
                             "[--dump-order FILE] [--dump-face-rank FILE] [--dump-rest FILE] [--eager] [--copy-calibration MB] "
                             "[--print-config] [--grasp R] [--cursor X,Y,Z] [--cursor-velocity VX,VY,VZ] "
                             "[--dump-attachments FILE] [--autotune] [--tune-cache FILE]\n", argv[0]);
        return 2;
    }
    std::string cfg_path = argv[1], dump_pos, dump_order, dump_rank, dump_rest, dump_att;
    double grasp_radius = -1;
    bool autotune = false;
    bool have_cursor = false;
    Sim::Vec3r cursor{ 0, 0, 0 }, cursor_vel{ 0, 0, 0 };
    auto triple = [](const char* t, Sim::Vec3r& out) {
        return std::sscanf(t, "%lf,%lf,%lf", &out[0], &out[1], &out[2]) == 3;
    };
    bool beam = false, print_only = false, load_meshes = false, eager = false;
    double copy_mb = 0;
    std::string solver_override, tune_cache;
    long steps = -1;
    for (int i = 2; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--beam") beam = true;
        else if (a == "--print-config") print_only = true;
        else if (a == "--load-meshes") load_meshes = true;
        else if (a == "--eager") eager = true;
        else if (a == "--autotune") autotune = true;
        else if (a == "--solver-type" && i + 1 < argc) solver_override = argv[++i];
        else if (a == "--tune-cache" && i + 1 < argc) tune_cache = argv[++i];
        else if (a == "--copy-calibration" && i + 1 < argc) copy_mb = std::atof(argv[++i]);
        else if (a == "--steps" && i + 1 < argc) steps = std::atol(argv[++i]);
        else if (a == "--dump-positions" && i + 1 < argc) dump_pos = argv[++i];
        else if (a == "--dump-order" && i + 1 < argc) dump_order = argv[++i];
        else if (a == "--dump-face-rank" && i + 1 < argc) dump_rank = argv[++i];
        else if (a == "--dump-rest" && i + 1 < argc) dump_rest = argv[++i];
        else if (a == "--dump-attachments" && i + 1 < argc) dump_att = argv[++i];
        else if (a == "--grasp" && i + 1 < argc) grasp_radius = std::atof(argv[++i]);
        else if (a == "--cursor" && i + 1 < argc && triple(argv[i + 1], cursor)) have_cursor = true, i++;
        else if (a == "--cursor-velocity" && i + 1 < argc && triple(argv[i + 1], cursor_vel)) i++;
        else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    try {
        Config::SimulationConfig cfg = Config::SimulationConfig::from_file(cfg_path);
        if (!solver_override.empty())
            for (auto& o : cfg.xpbd_objects) o.solver_type = Config::solver_type_from_string(solver_override);
        if (print_only) {
            std::printf("{\"name\": \"%s\", \"time_step\": %.9g, \"end_time\": %.9g, \"g_accel\": %.9g, "
                        "\"collision_rate\": %.9g, \"objects\": [",
                        cfg.name.c_str(), cfg.time_step, cfg.end_time, cfg.g_accel, cfg.collision_rate);
            for (size_t i = 0; i < cfg.xpbd_objects.size(); i++) {
                const auto& o = cfg.xpbd_objects[i];
                std::printf("%s{\"name\": \"%s\", \"solver_type\": \"%s\", \"iters\": %d, \"local_collision_iters\": %d, "
                        "\"E\": %.9g, \"nu\": %.9g, "
                            "\"density\": %.9g, \"collisions\": %s, \"source\": \"%s\", \"position\": [%.9g, %.9g, %.9g], "
                            "\"first_order\": %s, \"damping_multiplier\": %.9g, \"adjust_damping\": %s",
                            i ? ", " : "", o.name.c_str(), solver_name(o.solver_type), o.num_solver_iters,
                            o.num_local_collision_iters, o.material.E, o.material.nu, o.material.density, o.collisions ? "true" : "false",
                            o.grid ? "grid" : o.filename.c_str(), o.position[0], o.position[1], o.position[2],
                            o.first_order ? "true" : "false", o.damping_multiplier,
                            o.adjust_damping_to_material ? "true" : "false");
                std::printf(", \"materials\": [");
                for (size_t k = 0; k < o.materials.size(); k++)
                    std::printf("%s{\"name\": \"%s\", \"E\": %.9g, \"nu\": %.9g, \"density\": %.9g}", k ? ", " : "",
                                o.materials[k].name.c_str(), o.materials[k].E, o.materials[k].nu, o.materials[k].density);
                std::printf("], \"element_classes\": \"%s\"",
                            o.element_classes_filename ? o.element_classes_filename->c_str() : "");
                if (load_meshes) {
                    // host-side mesh set-up only (no device call)
                    Sim::XPBDMeshObject m(o, cfg.time_step, cfg.g_accel);
                    const Sim::AABB b = m.boundingBox();
                    m.descriptor();
                    std::printf(", \"vertices\": %d, \"elements\": %d, \"surface_faces\": %d, "
                                "\"bbox\": [[%.9g, %.9g, %.9g], [%.9g, %.9g, %.9g]]",
                                m.numVertices(), m.numElements(), m.numFaces(), b.min[0], b.min[1], b.min[2], b.max[0],
                                b.max[1], b.max[2]);
                    // elements per material after the class lookup, vertices per vertex class
                    std::vector<long> per_mat(m.config().materials.size(), 0), per_vclass;
                    for (int e = 0; e < m.numElements(); e++) per_mat[m.elementMaterialIndex(e)]++;
                    for (int v = 0; v < m.numVertices(); v++) {
                        const int c = std::max(0, m.vertexClass(v));
                        if ((size_t)c >= per_vclass.size()) per_vclass.resize(c + 1, 0);
                        per_vclass[c]++;
                    }
                    std::printf(", \"elements_per_material\": [");
                    for (size_t k = 0; k < per_mat.size(); k++) std::printf("%s%ld", k ? ", " : "", per_mat[k]);
                    std::printf("], \"vertices_per_class\": [");
                    for (size_t k = 0; k < per_vclass.size(); k++) std::printf("%s%ld", k ? ", " : "", per_vclass[k]);
                    std::printf("]");
                }
                std::printf("}");
            }
            std::printf("], \"rigid_objects\": %zu}\n", cfg.rigid_objects.size());
            return 0;
        }
        std::unique_ptr<Sim::Simulation> sim;
        Sim::GraspingSimulation* grasp = nullptr;
        if (beam) sim = std::make_unique<Sim::BeamSimulation>(cfg);
        else if (grasp_radius > 0) {
            auto g = std::make_unique<Sim::GraspingSimulation>(cfg);
            grasp = g.get();
            sim = std::move(g);
        } else
            sim = std::make_unique<Sim::Simulation>(cfg);
        if (copy_mb > 0) {
            const size_t bytes = (size_t)(copy_mb * 1048576.0);
            void *a = nullptr, *b = nullptr;
            if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
                hipMemset(a, 0, bytes) != hipSuccess || hipMemcpy(b, a, bytes, hipMemcpyDeviceToDevice) != hipSuccess ||
                hipDeviceSynchronize() != hipSuccess)
                throw Sim::SimError(XPBD_E_HIP, "calibration copy failed");
            (void)hipFree(a);
            (void)hipFree(b);
            std::fprintf(stderr, "calibration copy: %zu bytes device to device\n", bytes);
        }
        sim->setup();
        if (autotune)
            for (auto& o : sim->objects()) {
                const int rc = xpbd_autotune(o->solver().handle(), tune_cache.empty() ? nullptr : tune_cache.c_str());
                if (rc != XPBD_OK) throw Sim::SimError(rc, "xpbd_autotune");
                int32_t v = 0, b = 0, x = 0;
                xpbd_get_variant(o->solver().handle(), &v, &b, &x);
                std::fprintf(stderr, "autotuned %s: variant %d block %d xcd %d\n", o->name().c_str(), v, b, x);
            }
        if (grasp) {
            if (!have_cursor && !sim->objects().empty()) {
                const Sim::AABB b = sim->objects()[0]->boundingBox();
                cursor = { (b.min[0] + b.max[0]) / 2, (b.min[1] + b.max[1]) / 2, b.max[2] };
            }
            grasp->setGraspRadius(grasp_radius);
            grasp->setCursor(cursor);
            grasp->setCursorVelocity(cursor_vel);
            grasp->toggleGrasping();
            std::fprintf(stderr, "grasped %zu vertices\n", grasp->graspedVertices().size());
        }
        if (eager)
            for (auto& o : sim->objects()) {
                const int rc = xpbd_set_kernel_timing(o->solver().handle(), 1);
                if (rc != XPBD_OK) throw Sim::SimError(rc, "xpbd_set_kernel_timing");
            }
        sim->run(steps);
        std::cout << sim->logText();
        std::fprintf(stderr, "%lld steps, t = %.6g s\n", (long long)sim->steps(), sim->time());
        if (!sim->objects().empty()) {
            Sim::XPBDMeshObject& o = *sim->objects()[0];
            if (!dump_pos.empty()) {
                std::ofstream f(dump_pos, std::ios::binary);
                f.write(reinterpret_cast<const char*>(o.vertices().data()), sizeof(double) * o.vertices().size());
            }
            if (!dump_att.empty()) {
                std::vector<int32_t> v;
                if (grasp)
                    for (const auto& p : grasp->graspedVertices())
                        if (p.first == 0) v.push_back(p.second);
                std::ofstream f(dump_att, std::ios::binary);
                f.write(reinterpret_cast<const char*>(v.data()), sizeof(int32_t) * v.size());
            }
            if (!dump_rest.empty()) {
                std::ofstream f(dump_rest, std::ios::binary);
                f.write(reinterpret_cast<const char*>(o.restVertices().data()),
                        sizeof(double) * o.restVertices().size());
            }
            if (!dump_order.empty()) {
                std::vector<int32_t> order(o.numElements());
                const int rc = xpbd_get_tet_order(o.solver().handle(), order.data());
                if (rc != XPBD_OK) throw Sim::SimError(rc, "xpbd_get_tet_order");
                std::ofstream f(dump_order, std::ios::binary);
                f.write(reinterpret_cast<const char*>(order.data()), sizeof(int32_t) * order.size());
            }
            if (!dump_rank.empty()) {
                std::vector<int32_t> rank(o.numFaces());
                if (!rank.empty()) {
                    const int rc = xpbd_get_face_rank(o.solver().handle(), rank.data());
                    if (rc != XPBD_OK) throw Sim::SimError(rc, "xpbd_get_face_rank");
                }
                std::ofstream f(dump_rank, std::ios::binary);
                f.write(reinterpret_cast<const char*>(rank.data()), sizeof(int32_t) * rank.size());
            }
        }
    } catch (const Sim::SimError& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code, e.what());
        return 1;
    }
    return 0;
}
