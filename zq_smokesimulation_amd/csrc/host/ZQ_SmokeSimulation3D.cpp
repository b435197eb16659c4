// ZQ_SmokeSimulation3D (reference: ZQ_SmokeSimulation3D.cpp).  Two ways to run a frame:
//   host path    -- the reference's: each stage function moves the host grid through its C-ABI
//                   call (fields cross PCIe at every stage);
//   device path  -- deviceResident (default): the fields live in HBM in a zq3d_sim and the
//                   stages run back to back on one stream; the host grid is refreshed on demand.
// Solvers: TOTAL_{OPEN,CLOSED}_POISSON_SOR are the full-grid red-black SOR this library builds.
// COARSE_{OPEN,CLOSED}_POISSON_SOR (the reference's default) is the coarse + sub-grid projection
// (zq3d_config.sub_size = subSize); both run the open face-fraction solve on the coarse grid, as
// the reference's ZQ_TotalProjection3D does (ZQ_SmokeSimulationUtils3D.cpp:3730-3745).  The four
// *_FLUX_SOR solvers run the flux form of that solve (TOTAL_* on the fine grid, s = 1), and
// {OPEN,CLOSED}_OCTREE_POISSON_SOR the octree projection (zq3d_config.octree).  In the staged host
// path those projections are one device call on a projection-only simulation (the fields cross
// PCIe as in the other stages).  The CPU PCG solvers (plain and octree) are not built and Init
// fails for them.
#include "ZQ_SmokeSimulation3D.h"
#include "ZQ_SmokeGuideGrid3D.h"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

bool solver_built(GlobalSolverType3D s) {
    return s == TOTAL_OPEN_POISSON_SOR || s == TOTAL_CLOSED_POISSON_SOR || s == COARSE_OPEN_POISSON_SOR ||
           s == COARSE_CLOSED_POISSON_SOR || s == COARSE_OPEN_FLUX_SOR || s == COARSE_CLOSED_FLUX_SOR ||
           s == TOTAL_OPEN_FLUX_SOR || s == TOTAL_CLOSED_FLUX_SOR || s == OPEN_OCTREE_POISSON_SOR ||
           s == CLOSED_OCTREE_POISSON_SOR;
}

// counter-based hash noise, same as the device-resident test fixtures: lowbias32 of (index, seed)
inline uint32_t hash_u32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

void add_noise(float* T, size_t n, unsigned seed, float amp) {
    const uint32_t s = hash_u32(seed + 0x632BE5ABu);
    for (size_t o = 0; o < n; o++) {
        const uint32_t x = hash_u32((uint32_t)o * 0x9E3779B9u ^ s);
        const float u01 = (float)(x >> 8) * (1.0f / 16777216.0f);
        T[o] = T[o] + amp * (2.0f * u01 - 1.0f);
    }
}

}  // namespace

ZQ_SmokeSimulation3D::ZQ_SmokeSimulation3D()
    : fineGrid(nullptr), dev(nullptr), proj(nullptr), moving_frame(0), host_current(true), lastStats{} {}

ZQ_SmokeSimulation3D::~ZQ_SmokeSimulation3D() {
    if (dev) zq3d_destroy(dev);
    if (proj) zq3d_destroy(proj);
    delete fineGrid;
    for (ZQ_SmokeMovingObject3D* o : mvobjs) delete o;
}

bool ZQ_SmokeSimulation3D::Init(const char* file) {
    const double t0 = now_s();
    if (!para.LoadFromFile(file, true)) return false;
    if (!solver_built(para.globalGridSolver)) {
        std::printf("globalGridSolver %d is not built (the CPU PCG and octree PCG solvers): use a *_SOR solver\n",
                    (int)para.globalGridSolver);
        return false;
    }
    if (para.useMovingObject) {
        // the reference resolves the file relative to the working directory
        if (!LoadMovingObjects3D(para.movingObjectFile, mvobjs)) {
            std::printf("cannot load moving objects from %s\n", para.movingObjectFile);
            return false;
        }
    }
    if (!_initRegular()) return false;
    std::printf("init regular cost:%f\n", now_s() - t0);
    return true;
}

// ZQ_SmokeSimulation3D::_initRegular (ZQ_SmokeSimulation3D.cpp:293-394), fine grid only
bool ZQ_SmokeSimulation3D::_initRegular() {
    fineGrid = new ZQ_SmokeGrid3D(para.width, para.height, para.depth);
    if (!ZQ_InitFineGrid3D(fineGrid, para)) return false;
    if (para.useMovingObject && !ZQ_ApplyMovingObject3D(fineGrid, mvobjs, para)) return false;
    if (para.tempNoise != 0.0f)
        add_noise(fineGrid->GetTemperaturePtr(), (size_t)para.width * para.height * para.depth, para.noiseSeed,
                  para.tempNoise);
    // moving-object files and randomised reinjection (std::rand per cell, in the reference's cell
    // order) run on the host grid: those configs take the staged path
    if (para.deviceResident && !para.useMovingObject && !para.randDensity) {
        zq3d_config c;
        zq3d_config_default(&c, para.width, para.height, para.depth);
        c.voxel_size = para.voxelSize;
        c.step_time = para.stepTime;
        c.steps = para.steps;
        c.Tamb = para.Tamb;
        c.buoyancy = para.buoyancyCoeff;
        c.gravity = para.gravityCoeff;
        c.confine = para.confineCoeff;
        c.vel_atten = std::exp(-para.stepTime * para.velAttenCoeff);
        c.temp_atten = std::exp(-para.stepTime * para.tempAttenCoeff);
        c.dens_atten = std::exp(-para.stepTime * para.densityAttenCoeff);
        ZQ_CUDA_Advection3D::AdvectVelocityType vt;
        if (!map_vel_type(para.advVelType, &vt)) return false;
        c.adv_vel_type = (int)vt;
        switch (para.advScaType) {
        case ZQ_SmokeSimulationPara3D::ADV_SCA_MAC: c.adv_sca_type = ZQ_CUDA_Advection3D::ADV_SCA_MAC; break;
        case ZQ_SmokeSimulationPara3D::ADV_SCA_MAC_BFECC: c.adv_sca_type = ZQ_CUDA_Advection3D::ADV_SCA_MAC_BFECC; break;
        case ZQ_SmokeSimulationPara3D::ADV_SCA_REG: c.adv_sca_type = ZQ_CUDA_Advection3D::ADV_SCA_REG; break;
        case ZQ_SmokeSimulationPara3D::ADV_SCA_REG_BFECC: c.adv_sca_type = ZQ_CUDA_Advection3D::ADV_SCA_REG_BFECC; break;
        }
        const bool closed = para.IsClosedSolver();
        c.sor.open_mask = para.OpenFaceMask();
        c.sor.closed_mean = closed ? 1 : 0;
        c.sor.omega = para.omega;
        c.sor.iters = para.maxIter;
        c.sor.mixed_precision = para.mixedPrecision ? 1 : 0;
        c.sor.max_outer = para.maxOuter;
        c.sor.tol = para.sorTol;
        c.warm_start = para.warmStart ? 1 : 0;
        if (device_projection()) set_projection(&c);
        if (zq3d_create(&c, &dev) != ZQ3D_OK) {
            std::printf("device simulation: %s\n", zq3d_last_error());
            return false;
        }
        std::vector<float> boxes, spheres;
        for (const ZQ_SmokeSource3D& s : para.sources)
            boxes.insert(boxes.end(), {s.cx, s.cy, s.cz, s.half_xlen, s.half_ylen, s.half_zlen, s.reinjectDensity,
                                       s.reinjectTemperature});
        for (const ZQ_SmokeSphereSource3D& s : para.sphereSources)
            spheres.insert(spheres.end(), {s.cx, s.cy, s.cz, s.radius, s.reinjectDensity, s.reinjectTemperature});
        zq3d_set_box_sources(dev, (int)boxes.size() / 8, boxes.data());
        zq3d_set_sphere_sources(dev, (int)spheres.size() / 6, spheres.data());
        const size_t N = (size_t)para.width * para.height * para.depth;
        std::vector<uint8_t> occ(N);
        for (size_t o = 0; o < N; o++) occ[o] = fineGrid->GetOccupyPtr()[o] ? 1 : 0;
        const size_t Nu = (size_t)(para.width + 1) * para.height * para.depth;
        const size_t Nv = (size_t)para.width * (para.height + 1) * para.depth;
        const size_t Nw = (size_t)para.width * para.height * (para.depth + 1);
        if (zq3d_upload(dev, ZQ3D_FIELD_OCC, occ.data(), N) || zq3d_upload(dev, ZQ3D_FIELD_U, fineGrid->GetVelocityUptr(), Nu * 4) ||
            zq3d_upload(dev, ZQ3D_FIELD_V, fineGrid->GetVelocityVptr(), Nv * 4) ||
            zq3d_upload(dev, ZQ3D_FIELD_W, fineGrid->GetVelocityWptr(), Nw * 4) ||
            zq3d_upload(dev, ZQ3D_FIELD_T, fineGrid->GetTemperaturePtr(), N * 4) ||
            zq3d_upload(dev, ZQ3D_FIELD_RHO, fineGrid->GetDensityPtr(), N * 4)) {
            std::printf("device upload: %s\n", zq3d_last_error());
            return false;
        }
    }
    if (!dev && device_projection()) {
        // projection-only device simulation of the staged path
        zq3d_config c;
        zq3d_config_default(&c, para.width, para.height, para.depth);
        c.sor.omega = para.omega;
        set_projection(&c);
        if (zq3d_create(&c, &proj) != ZQ3D_OK) {
            std::printf("device projection: %s\n", zq3d_last_error());
            return false;
        }
    }
    host_current = true;
    return true;
}

bool ZQ_SmokeSimulation3D::is_coarse() const {
    zq3d_config c{};
    return map_coarse_solver(para, &c);
}

bool ZQ_SmokeSimulation3D::device_projection() const {
    zq3d_config c{};
    return map_coarse_solver(para, &c) || map_octree_solver(para, &c);
}

// the coarse-grid solves (COARSE_* and the flux solvers) and the octree solves
void ZQ_SmokeSimulation3D::set_projection(zq3d_config* c) const {
    if (!map_coarse_solver(para, c)) map_octree_solver(para, c);
}

bool ZQ_SmokeSimulation3D::SyncToHost() {
    if (!dev || host_current) return true;
    const size_t N = (size_t)para.width * para.height * para.depth;
    const size_t Nu = (size_t)(para.width + 1) * para.height * para.depth;
    const size_t Nv = (size_t)para.width * (para.height + 1) * para.depth;
    const size_t Nw = (size_t)para.width * para.height * (para.depth + 1);
    std::vector<uint8_t> occ(N);
    if (zq3d_download(dev, ZQ3D_FIELD_U, fineGrid->GetVelocityUptr(), Nu * 4) ||
        zq3d_download(dev, ZQ3D_FIELD_V, fineGrid->GetVelocityVptr(), Nv * 4) ||
        zq3d_download(dev, ZQ3D_FIELD_W, fineGrid->GetVelocityWptr(), Nw * 4) ||
        zq3d_download(dev, ZQ3D_FIELD_T, fineGrid->GetTemperaturePtr(), N * 4) ||
        zq3d_download(dev, ZQ3D_FIELD_RHO, fineGrid->GetDensityPtr(), N * 4) ||
        zq3d_download(dev, ZQ3D_FIELD_OCC, occ.data(), N)) {
        std::printf("device download: %s\n", zq3d_last_error());
        return false;
    }
    for (size_t o = 0; o < N; o++) fineGrid->GetOccupyPtr()[o] = occ[o] != 0;
    host_current = true;
    return true;
}

bool ZQ_SmokeSimulation3D::Run(const char* fold) {
    char name[1024];
    // one guide file per frame, <guidingFold>/<frame>.guidedat at the coarse resolution
    // (ZQ_SmokeSimulation3D.cpp:110-124; the reference joins the path with a backslash)
    ZQ_SmokeGuideGrid3D guideGrid(para.coarseWidth, para.coarseHeight, para.coarseDepth);
    for (int i = 0; i < para.frameCount; i++) {
        std::printf("frame [%d]\n", i);
        if (para.useGuidingControl) {
            std::snprintf(name, sizeof(name), "%s/%d.guidedat", para.guidingFold, i);
            if (!guideGrid.LoadFromFile(name)) {
                std::printf("load guide file %s fail\n", name);
                return false;
            }
        }
        if (!UpdateOneFrame(para.useGuidingControl ? &guideGrid : nullptr)) return false;
        if (!SyncToHost()) return false;
        switch (para.exportType) {
        case ZQ_SmokeSimulationPara3D::EXPORT_DI3: std::snprintf(name, sizeof(name), "%s/frame%d.di3", fold, i); break;
        case ZQ_SmokeSimulationPara3D::EXPORT_ZQCI: std::snprintf(name, sizeof(name), "%s/frame%d.zqci", fold, i); break;
        case ZQ_SmokeSimulationPara3D::EXPORT_ZQWV: std::snprintf(name, sizeof(name), "%s/frame%d.zqwv", fold, i); break;
        }
        if (!fineGrid->ExportDensity(name, para.export_quality)) std::printf("export %s failed\n", name);
        if (para.exportGuide) {
            std::snprintf(name, sizeof(name), "%s/%d.guidedat", fold, i);
            fineGrid->ExportGuide(name);
        }
        if (para.exportOccupy) {
            std::snprintf(name, sizeof(name), "%s/%d.occdat", fold, i);
            fineGrid->ExportOccupy(name);
        }
    }
    return true;
}

bool ZQ_SmokeSimulation3D::UpdateOneFrame(const ZQ_SmokeGuideGrid3D* guideGrid) {
    moving_frame += 1;
    // guiding acts in the coarse projection and, as the special-region step, in the octree
    // projection (ZQ_SmokeSimulation3D.cpp:500-508); the reference ignores it for the full-grid
    // solvers (:403-414)
    zq3d_sim* target = dev ? dev : proj;
    zq3d_config oc{};
    if (target && map_octree_solver(para, &oc)) {
        const bool on = guideGrid && para.useGuidingControl;
        const float unread = 0.0f;  // any non-null pointer: the octree step reads no guide values
        if (zq3d_set_guide(target, para.guidingCoeff, on ? &unread : nullptr, nullptr, nullptr) != ZQ3D_OK) {
            std::printf("guide: %s\n", zq3d_last_error());
            return false;
        }
    } else if (target && is_coarse()) {
        const bool on = guideGrid && para.useGuidingControl;
        if (on && ((int)guideGrid->GetWidth() != para.coarseWidth || (int)guideGrid->GetHeight() != para.coarseHeight ||
                   (int)guideGrid->GetDepth() != para.coarseDepth)) {
            std::printf("guide grid is not at the coarse resolution\n");
            return false;
        }
        if (zq3d_set_guide(target, para.guidingCoeff, on ? guideGrid->GetUptr() : nullptr,
                           on ? guideGrid->GetVptr() : nullptr, on ? guideGrid->GetWptr() : nullptr) != ZQ3D_OK) {
            std::printf("guide: %s\n", zq3d_last_error());
            return false;
        }
    }
    return dev ? _updateResident() : _updateHost();
}

// the reference's frame (ZQ_SmokeSimulation3D.cpp:156-243), stage by stage through the C ABI
bool ZQ_SmokeSimulation3D::_updateHost() {
    double t = now_s();
    if (para.useMovingObject && moving_frame > 1 && !ZQ_ApplyMovingObject3D(fineGrid, mvobjs, para)) return false;
    std::printf("app boundary cost:%f\n", now_s() - t);
    t = now_s();
    if (!ZQ_AddForce3D(fineGrid, para)) return false;
    std::printf("add force cost:%f\n", now_s() - t);
    t = now_s();
    if (!ZQ_Attenuation3D(fineGrid, para)) return false;
    std::printf("attenuation cost:%f\n", now_s() - t);
    t = now_s();
    if (!ZQ_AdvectVelocity3D(fineGrid, para)) return false;
    std::printf("advect velocity cost:%f\n", now_s() - t);
    t = now_s();
    if (!_projection()) return false;
    std::printf("solve global cost:%f\n", now_s() - t);
    t = now_s();
    if (!ZQ_AdvectScalar3D(fineGrid, para)) return false;
    std::printf("advect tempden cost:%f\n", now_s() - t);
    if (!ZQ_ReinjectTemperatureDensity3D(fineGrid, para)) return false;
    if (para.useMovingObject)
        for (ZQ_SmokeMovingObject3D* o : mvobjs)
            if (o) o->UpdateOneStep();
    return true;
}

// _projection / _totalProjection (ZQ_SmokeSimulation3D.cpp:396-549) for the built solvers
bool ZQ_SmokeSimulation3D::_projection() {
    if (proj) {
        const size_t N = (size_t)para.width * para.height * para.depth;
        const size_t Nu = (size_t)(para.width + 1) * para.height * para.depth;
        const size_t Nv = (size_t)para.width * (para.height + 1) * para.depth;
        const size_t Nw = (size_t)para.width * para.height * (para.depth + 1);
        std::vector<uint8_t> occ(N);
        for (size_t o = 0; o < N; o++) occ[o] = fineGrid->GetOccupyPtr()[o] ? 1 : 0;
        if (zq3d_upload(proj, ZQ3D_FIELD_OCC, occ.data(), N) ||
            zq3d_upload(proj, ZQ3D_FIELD_U, fineGrid->GetVelocityUptr(), Nu * 4) ||
            zq3d_upload(proj, ZQ3D_FIELD_V, fineGrid->GetVelocityVptr(), Nv * 4) ||
            zq3d_upload(proj, ZQ3D_FIELD_W, fineGrid->GetVelocityWptr(), Nw * 4) || zq3d_project(proj, &lastStats) ||
            zq3d_download(proj, ZQ3D_FIELD_U, fineGrid->GetVelocityUptr(), Nu * 4) ||
            zq3d_download(proj, ZQ3D_FIELD_V, fineGrid->GetVelocityVptr(), Nv * 4) ||
            zq3d_download(proj, ZQ3D_FIELD_W, fineGrid->GetVelocityWptr(), Nw * 4)) {
            std::printf("coarse projection: %s\n", zq3d_last_error());
            return false;
        }
        return true;
    }
    if (para.IsClosedSolver()) return ZQ_SolveClosedPressure3DSOR(fineGrid, para);
    return ZQ_SolveOpenPressure3DSOR(fineGrid, para);
}

bool ZQ_SmokeSimulation3D::_updateResident() {
    const double t = now_s();
    if (zq3d_step(dev, nullptr, &lastStats) != ZQ3D_OK) {
        std::printf("step failed: %s\n", zq3d_last_error());
        return false;
    }
    host_current = false;
    std::printf("step cost:%f (projection: %d sweeps, max|div| %g)\n", now_s() - t, lastStats.sweeps,
                lastStats.max_div);
    return true;
}
