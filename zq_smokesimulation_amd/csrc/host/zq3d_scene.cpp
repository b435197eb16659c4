// Host-only C ABI (Part 4 of zq_smoke3d.h): the reference's config parser and grid set-up,
// usable without a GPU.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "ZQ_SmokeMovingObject3D.h"
#include "ZQ_SmokeSimulationUtils3D.h"

extern "C" int zq3d_scene_load(const char* path, zq3d_scene* out) {
    if (!path || !out) return ZQ3D_ERR_ARG;
    ZQ_SmokeSimulationPara3D p;
    if (!p.LoadFromFile(path, false)) return ZQ3D_ERR_ARG;
    std::memset(out, 0, sizeof(*out));
    out->struct_size = (int)sizeof(zq3d_scene);
    zq3d_config& c = out->config;
    zq3d_config_default(&c, p.width, p.height, p.depth);
    c.voxel_size = p.voxelSize;
    c.step_time = p.stepTime;
    c.steps = p.steps;
    c.Tamb = p.Tamb;
    c.buoyancy = p.buoyancyCoeff;
    c.gravity = p.gravityCoeff;
    c.confine = p.confineCoeff;
    c.vel_atten = std::exp(-p.stepTime * p.velAttenCoeff);
    c.temp_atten = std::exp(-p.stepTime * p.tempAttenCoeff);
    c.dens_atten = std::exp(-p.stepTime * p.densityAttenCoeff);
    static const int vel_map[] = {ZQ_CUDA_Advection3D::ADV_VEL_INMAC_OUTMAC, ZQ_CUDA_Advection3D::ADV_VEL_INMAC_OUTMAC_BFECC,
                                  ZQ_CUDA_Advection3D::ADV_VEL_INREG_OUTREG, ZQ_CUDA_Advection3D::ADV_VEL_INREG_OUTREG_BFECC,
                                  ZQ_CUDA_Advection3D::ADV_VEL_INREG_OUTMAC, ZQ_CUDA_Advection3D::ADV_VEL_INREG_OUTMAC_BFECC};
    c.adv_vel_type = vel_map[p.advVelType];
    c.adv_sca_type = (int)p.advScaType;  // both enums list MAC, MAC_BFECC, Reg, Reg_BFECC in this order
    c.sor.open_mask = p.OpenFaceMask();
    c.sor.closed_mean = p.IsClosedSolver() ? 1 : 0;
    c.sor.omega = p.omega;
    c.sor.iters = p.maxIter;
    c.sor.mixed_precision = p.mixedPrecision ? 1 : 0;
    c.sor.max_outer = p.maxOuter;
    c.sor.tol = p.sorTol;
    c.warm_start = p.warmStart ? 1 : 0;
    map_coarse_solver(p, &c);
    map_octree_solver(p, &c);
    out->frame_count = p.frameCount;
    out->solver = (int)p.globalGridSolver;
    out->wall_faces = p.WallFaceMask();
    out->internal_object = (int)p.intobj;
    out->use_moving_object = p.useMovingObject ? 1 : 0;
    out->export_type = (int)p.exportType;
    out->device_resident = p.deviceResident ? 1 : 0;
    out->sub_size = p.subSize;
    out->max_iter = p.maxIter;
    out->inner_iter = p.innerIter;
    out->steps = p.steps;
    out->vel_atten_coeff = p.velAttenCoeff;
    out->temp_atten_coeff = p.tempAttenCoeff;
    out->dens_atten_coeff = p.densityAttenCoeff;
    out->adv_vel_type_para = (int)p.advVelType;
    out->adv_sca_type_para = (int)p.advScaType;
    out->add_force_type = (int)p.addforce_type;
    out->n_box_sources = (int)std::min<size_t>(p.sources.size(), ZQ3D_MAX_SOURCES);
    for (int i = 0; i < out->n_box_sources; i++) {
        const ZQ_SmokeSource3D& s = p.sources[i];
        const float v[8] = {s.cx, s.cy, s.cz, s.half_xlen, s.half_ylen, s.half_zlen, s.reinjectDensity,
                            s.reinjectTemperature};
        std::memcpy(out->box_sources[i], v, sizeof(v));
    }
    out->n_sphere_sources = (int)std::min<size_t>(p.sphereSources.size(), ZQ3D_MAX_SOURCES);
    for (int i = 0; i < out->n_sphere_sources; i++) {
        const ZQ_SmokeSphereSource3D& s = p.sphereSources[i];
        const float v[6] = {s.cx, s.cy, s.cz, s.radius, s.reinjectDensity, s.reinjectTemperature};
        std::memcpy(out->sphere_sources[i], v, sizeof(v));
    }
    out->temp_noise = p.tempNoise;
    out->noise_seed = p.noiseSeed;
    std::memcpy(out->moving_object_file, p.movingObjectFile, sizeof(out->moving_object_file) - 1);
    out->moving_object_file[sizeof(out->moving_object_file) - 1] = '\0';
    return ZQ3D_OK;
}

extern "C" int zq3d_scene_init_occupancy(const char* path, uint8_t* occ, size_t bytes) {
    if (!path || !occ) return ZQ3D_ERR_ARG;
    ZQ_SmokeSimulationPara3D p;
    if (!p.LoadFromFile(path, false)) return ZQ3D_ERR_ARG;
    const size_t N = (size_t)p.width * p.height * p.depth;
    if (bytes != N) return ZQ3D_ERR_ARG;
    ZQ_SmokeGrid3D g(p.width, p.height, p.depth);
    if (!ZQ_InitFineGrid3D(&g, p)) return ZQ3D_ERR_ARG;
    for (size_t o = 0; o < N; o++) occ[o] = g.GetOccupyPtr()[o] ? 1 : 0;
    return ZQ3D_OK;
}

extern "C" int zq3d_scene_apply_moving_objects(const char* path, int frames, uint8_t* occ, float* u, float* v,
                                               float* w) {
    if (!path || !occ || !u || !v || !w || frames < 0) return ZQ3D_ERR_ARG;
    ZQ_SmokeSimulationPara3D p;
    if (!p.LoadFromFile(path, false)) return ZQ3D_ERR_ARG;
    std::vector<ZQ_SmokeMovingObject3D*> objs;
    int rc = ZQ3D_OK;
    if (!LoadMovingObjects3D(p.movingObjectFile, objs)) rc = ZQ3D_ERR_ARG;
    if (rc == ZQ3D_OK) {
        const size_t N = (size_t)p.width * p.height * p.depth;
        ZQ_SmokeGrid3D g(p.width, p.height, p.depth);
        std::memcpy(g.GetVelocityUptr(), u, sizeof(float) * (p.width + 1) * p.height * p.depth);
        std::memcpy(g.GetVelocityVptr(), v, sizeof(float) * p.width * (p.height + 1) * p.depth);
        std::memcpy(g.GetVelocityWptr(), w, sizeof(float) * p.width * p.height * (p.depth + 1));
        for (int f = 0; f < frames; f++)
            for (ZQ_SmokeMovingObject3D* o : objs) o->UpdateOneStep();
        ZQ_ApplyMovingObject3D(&g, objs, p);
        for (size_t o = 0; o < N; o++) occ[o] = g.GetOccupyPtr()[o] ? 1 : 0;
        std::memcpy(u, g.GetVelocityUptr(), sizeof(float) * (p.width + 1) * p.height * p.depth);
        std::memcpy(v, g.GetVelocityVptr(), sizeof(float) * p.width * (p.height + 1) * p.depth);
        std::memcpy(w, g.GetVelocityWptr(), sizeof(float) * p.width * p.height * (p.depth + 1));
    }
    for (ZQ_SmokeMovingObject3D* o : objs) delete o;
    return rc;
}

// .di3 (the density container ZQ_SmokeGrid3D::ExportDensity writes, ZQ_SmokeGrid3D.cpp:55-81): a
// 16-byte type tag holding a NUL-terminated name ("float"; bytes after the NUL are not part of
// it), int32 width, height, depth, channels, one flag byte, then width*height*depth*channels FP32
// samples, x fastest.
extern "C" int zq3d_di3_read(const char* path, int* dims4, float* data, size_t capacity) {
    if (!path || !dims4) return ZQ3D_ERR_ARG;
    FILE* f = std::fopen(path, "rb");
    if (!f) return ZQ3D_ERR_ARG;
    char tag[16];
    int hdr[4];
    unsigned char flag = 0;
    int rc = ZQ3D_OK;
    if (std::fread(tag, 1, 16, f) != 16 || std::fread(hdr, 4, 4, f) != 4 || std::fread(&flag, 1, 1, f) != 1 ||
        std::strncmp(tag, "float", 6) != 0 || hdr[0] < 1 || hdr[1] < 1 || hdr[2] < 1 || hdr[3] < 1) {
        rc = ZQ3D_ERR_ARG;
    } else {
        std::memcpy(dims4, hdr, sizeof(hdr));
        const size_t n = (size_t)hdr[0] * hdr[1] * hdr[2] * hdr[3];
        if (data) {
            if (capacity < n || std::fread(data, 4, n, f) != n) rc = ZQ3D_ERR_ARG;
        } else {
            // header only: the payload must still be whole
            if (std::fseek(f, 0, SEEK_END) != 0 || std::ftell(f) != (long)(33 + 4 * n)) rc = ZQ3D_ERR_ARG;
        }
    }
    std::fclose(f);
    return rc;
}
