// ZQ_SmokeSimulationUtils3D -- host side of the smoke step: parameters, the reference's config
// format, and the stage functions.  Every stage calls the C ABI exactly as the reference's stage
// functions call their CUDA wrappers (ZQ_SmokeSimulationUtils3D.cpp:2049-2168, 2349-2440, 3796-3882).
#include "ZQ_SmokeSimulationUtils3D.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <strings.h>

// defaults of ZQ_SmokeSimulationPara3D::ZQ_SmokeSimulationPara3D (ZQ_SmokeSimulationUtils3D.cpp:18-72)
ZQ_SmokeSimulationPara3D::ZQ_SmokeSimulationPara3D() {
    width = height = depth = 128;
    subSize = 4;
    coarseWidth = width / subSize;
    coarseHeight = height / subSize;
    coarseDepth = depth / subSize;
    gravityCoeff = 0;
    buoyancyCoeff = 1;
    confineCoeff = 0;
    Tamb = 10;
    voxelSize = (float)(1.0 / width);
    ZQ_SmokeSource3D source;
    source.cx = 0.5f;
    source.cy = 0.1f;
    source.cz = 0.5f;
    source.half_xlen = 0.1f;
    source.half_ylen = 0.05f;
    source.half_zlen = 0.1f;
    source.reinjectDensity = 1;
    source.reinjectTemperature = Tamb * 20;
    sources.push_back(source);
    velAttenCoeff = 0;
    densityAttenCoeff = 0;
    tempAttenCoeff = 0;
    stepTime = 0.01f;
    steps = 20;
    frameCount = 100;
    maxIter = 100;
    innerIter = 30;
    globalGridSolver = COARSE_OPEN_POISSON_SOR;
    subGridSolver = CLOSED_POISSON;
    octree_thresh1 = 2;
    octree_thresh2 = 4;
    octree_thresh3 = 16;
    useMovingObject = false;
    std::strcpy(movingObjectFile, "");
    useGuidingControl = false;
    guidingCoeff = 0;
    guidingNaiveBlend = true;
    exportGuide = false;
    std::strcpy(guidingFold, "");
    exportOccupy = false;
    randDensity = false;
    maxDensity = 0.8f;
    maxTemperature = 20.0f;
    intobj = INT_OBJ_NULL;
    advVelType = ADV_VEL_INREG_OUTREG;
    advScaType = ADV_SCA_REG;
    exportType = EXPORT_DI3;
    export_quality = 0.9999f;
    addforce_type = ADD_FORCE_ENTIRE;

    omega = 1.0f;
    mixedPrecision = false;
    sorTol = 1e-6f;
    maxOuter = 1000;
    openFaces = -1;
    wallFaces = -1;
    tempNoise = 0.0f;
    noiseSeed = 0;
    deviceResident = true;
    warmStart = true;
}

namespace {

bool eq(const char* a, const char* b) { return strcasecmp(a, b) == 0; }
bool truth(const char* a) { return strcasecmp(a, "true") == 0; }

// "63", "0x3f", or names joined by '|' / ',': XLO XHI YLO YHI ZLO ZHI ALL NONE
int parse_faces(const char* s) {
    char* end = nullptr;
    const long v = std::strtol(s, &end, 0);
    if (end && end != s && *end == '\0') return (int)(v & 63);
    int mask = 0;
    char buf[128];
    std::strncpy(buf, s, sizeof(buf) - 1);
    buf[sizeof(buf) - 1] = '\0';
    for (char* tok = std::strtok(buf, "|,+"); tok; tok = std::strtok(nullptr, "|,+")) {
        static const char* names[] = {"XLO", "XHI", "YLO", "YHI", "ZLO", "ZHI"};
        for (int i = 0; i < 6; i++)
            if (eq(tok, names[i])) mask |= 1 << i;
        if (eq(tok, "ALL")) mask = 63;
    }
    return mask;
}

}  // namespace

// ZQ_SmokeSimulationPara3D::LoadFromFile (ZQ_SmokeSimulationUtils3D.cpp:78-352).  Keeps the
// reference's behaviour: keys are case-insensitive, unknown keys are ignored, an unknown
// advVelType/advScaType/exportType/AddForceType value fails, an unknown globalGridSolver becomes
// OPEN_FLUX, "exportGuide" compares its value case-sensitively, "step" is not "steps", the
// voxel size is 1/width, and the grid must be divisible by subSize.
bool ZQ_SmokeSimulationPara3D::LoadFromFile(const char* file, bool display) {
    FILE* config = std::fopen(file, "r");
    if (!config) return false;
    sources.clear();
    sphereSources.clear();
    char s[1000];
    for (;;) {
        if (!std::fgets(s, 999, config)) break;
        char arg1[100] = {0}, arg2[100] = {0};
        if (std::sscanf(s, "%99s%99s", arg1, arg2) < 1) continue;
        if (eq(arg1, "resolution")) std::sscanf(s, "%*s%d%d%d", &width, &height, &depth);
        else if (eq(arg1, "subSize")) subSize = std::atoi(arg2);
        else if (eq(arg1, "Tamb")) Tamb = (float)std::atof(arg2);
        else if (eq(arg1, "velAttenCoeff")) velAttenCoeff = (float)std::atof(arg2);
        else if (eq(arg1, "tempAttenCoeff")) tempAttenCoeff = (float)std::atof(arg2);
        else if (eq(arg1, "densityAttenCoeff")) densityAttenCoeff = (float)std::atof(arg2);
        else if (eq(arg1, "confineCoeff")) confineCoeff = (float)std::atof(arg2);
        else if (eq(arg1, "source")) {
            ZQ_SmokeSource3D src{};
            std::sscanf(s, "%*s%f%f%f%f%f%f%f%f", &src.cx, &src.cy, &src.cz, &src.half_xlen, &src.half_ylen,
                        &src.half_zlen, &src.reinjectDensity, &src.reinjectTemperature);
            sources.push_back(src);
        } else if (eq(arg1, "steps")) steps = std::atoi(arg2);
        else if (eq(arg1, "stepTime")) stepTime = (float)std::atof(arg2);
        else if (eq(arg1, "frameCount")) frameCount = std::atoi(arg2);
        else if (eq(arg1, "maxIter")) maxIter = std::atoi(arg2);
        else if (eq(arg1, "innerIter")) innerIter = std::atoi(arg2);
        else if (eq(arg1, "globalGridSolver")) {
            static const struct { const char* name; GlobalSolverType3D v; } tab[] = {
                {"CLOSED_POISSON", CLOSED_POISSON}, {"OPEN_POISSON", OPEN_POISSON}, {"CLOSED_FLUX", CLOSED_FLUX},
                {"OPEN_FLUX", OPEN_FLUX}, {"CLOSED_OCTREE_FLUX", CLOSED_OCTREE_FLUX},
                {"OPEN_OCTREE_FLUX", OPEN_OCTREE_FLUX}, {"CLOSED_OCTREE_POISSON", CLOSED_OCTREE_POISSON},
                {"OPEN_OCTREE_POISSON", OPEN_OCTREE_POISSON}, {"TOTAL_OPEN_POISSON_SOR", TOTAL_OPEN_POISSON_SOR},
                {"TOTAL_CLOSED_POISSON_SOR", TOTAL_CLOSED_POISSON_SOR}, {"TOTAL_OPEN_FLUX_SOR", TOTAL_OPEN_FLUX_SOR},
                {"TOTAL_CLOSED_FLUX_SOR", TOTAL_CLOSED_FLUX_SOR}, {"COARSE_OPEN_POISSON_SOR", COARSE_OPEN_POISSON_SOR},
                {"COARSE_CLOSED_POISSON_SOR", COARSE_CLOSED_POISSON_SOR},
                {"COARSE_OPEN_FLUX_SOR", COARSE_OPEN_FLUX_SOR}, {"COARSE_CLOSED_FLUX_SOR", COARSE_CLOSED_FLUX_SOR},
                {"OPEN_OCTREE_POISSON_SOR", OPEN_OCTREE_POISSON_SOR},
                {"CLOSED_OCTREE_POISSON_SOR", CLOSED_OCTREE_POISSON_SOR}};
            globalGridSolver = OPEN_FLUX;
            for (const auto& t : tab)
                if (eq(arg2, t.name)) globalGridSolver = t.v;
        } else if (eq(arg1, "octree_thresh1")) octree_thresh1 = std::atoi(arg2);
        else if (eq(arg1, "octree_thresh2")) octree_thresh2 = std::atoi(arg2);
        else if (eq(arg1, "octree_thresh3")) octree_thresh3 = std::atoi(arg2);
        else if (eq(arg1, "useMovingObject")) useMovingObject = truth(arg2);
        else if (eq(arg1, "movingObjectFile")) std::strcpy(movingObjectFile, arg2);
        else if (eq(arg1, "useGuidingControl")) useGuidingControl = truth(arg2);
        else if (eq(arg1, "guidingCoeff")) guidingCoeff = (float)std::atof(arg2);
        else if (eq(arg1, "guidingNaiveBlend")) guidingNaiveBlend = truth(arg2);
        else if (eq(arg1, "guidingFold")) std::strcpy(guidingFold, arg2);
        else if (eq(arg1, "exportGuide")) exportGuide = std::strcmp(arg2, "true") == 0;
        else if (eq(arg1, "exportOccupy")) exportOccupy = truth(arg2);
        else if (eq(arg1, "randDensity")) randDensity = truth(arg2);
        else if (eq(arg1, "intobj")) {
            if (eq(arg2, "sphere")) intobj = INT_OBJ_SPHERE;
            else if (eq(arg2, "oval")) intobj = INT_OBJ_OVAL;
            else if (eq(arg2, "box")) intobj = INT_OBJ_BOX;
            else if (eq(arg2, "special1")) intobj = INT_OBJ_SPECIAL1;
            else intobj = INT_OBJ_NULL;
        } else if (eq(arg1, "advVelType")) {
            if (eq(arg2, "inMAC_outMAC")) advVelType = ADV_VEL_INMAC_OUTMAC;
            else if (eq(arg2, "inMAC_outMAC_BFECC")) advVelType = ADV_VEL_INMAC_OUTMAC_BFECC;
            else if (eq(arg2, "inReg_outReg")) advVelType = ADV_VEL_INREG_OUTREG;
            else if (eq(arg2, "inReg_outReg_BFECC")) advVelType = ADV_VEL_INREG_OUTREG_BFECC;
            else if (eq(arg2, "inReg_outMAC")) advVelType = ADV_VEL_INREG_OUTMAC;
            else if (eq(arg2, "inReg_outMAC_BFECC")) advVelType = ADV_VEL_INREG_OUTMAC_BFECC;
            else {
                std::printf("unknown para: %s: %s\n", arg1, arg2);
                std::fclose(config);
                return false;
            }
        } else if (eq(arg1, "advScaType")) {
            if (eq(arg2, "MAC")) advScaType = ADV_SCA_MAC;
            else if (eq(arg2, "MAC_BFECC")) advScaType = ADV_SCA_MAC_BFECC;
            else if (eq(arg2, "Reg")) advScaType = ADV_SCA_REG;
            else if (eq(arg2, "Reg_BFECC")) advScaType = ADV_SCA_REG_BFECC;
            else {
                std::printf("unknown para: %s: %s\n", arg1, arg2);
                std::fclose(config);
                return false;
            }
        } else if (eq(arg1, "exportType")) {
            if (eq(arg2, "DI3")) exportType = EXPORT_DI3;
            else if (eq(arg2, "ZQWV")) exportType = EXPORT_ZQWV;
            else if (eq(arg2, "ZQCI")) exportType = EXPORT_ZQCI;
            else {
                std::printf("unknown para: %s: %s\n", arg1, arg2);
                std::fclose(config);
                return false;
            }
        } else if (eq(arg1, "exportQuality")) export_quality = (float)std::atof(arg2);
        else if (eq(arg1, "AddForceType")) {
            if (eq(arg2, "ENTIRE")) addforce_type = ADD_FORCE_ENTIRE;
            else if (eq(arg2, "SLICE")) addforce_type = ADD_FORCE_SLICE;
            else {
                std::printf("unknown para: %s: %s\n", arg1, arg2);
                std::fclose(config);
                return false;
            }
        } else if (eq(arg1, "maxDensity")) maxDensity = (float)std::atof(arg2);
        else if (eq(arg1, "maxTemperature")) maxTemperature = (float)std::atof(arg2);
        // ---- extension keys ----
        else if (eq(arg1, "buoyancyCoeff")) buoyancyCoeff = (float)std::atof(arg2);
        else if (eq(arg1, "gravityCoeff")) gravityCoeff = (float)std::atof(arg2);
        else if (eq(arg1, "sphereSource")) {
            ZQ_SmokeSphereSource3D src{};
            std::sscanf(s, "%*s%f%f%f%f%f%f", &src.cx, &src.cy, &src.cz, &src.radius, &src.reinjectDensity,
                        &src.reinjectTemperature);
            sphereSources.push_back(src);
        } else if (eq(arg1, "omega")) omega = (float)std::atof(arg2);
        else if (eq(arg1, "mixedPrecision")) mixedPrecision = truth(arg2);
        else if (eq(arg1, "sorTol")) sorTol = (float)std::atof(arg2);
        else if (eq(arg1, "maxOuter")) maxOuter = std::atoi(arg2);
        else if (eq(arg1, "openFaces")) openFaces = parse_faces(arg2);
        else if (eq(arg1, "wallFaces")) wallFaces = parse_faces(arg2);
        else if (eq(arg1, "tempNoise")) {
            unsigned seed = 0;
            std::sscanf(s, "%*s%f%u", &tempNoise, &seed);
            noiseSeed = seed;
        } else if (eq(arg1, "deviceResident")) deviceResident = truth(arg2);
        else if (eq(arg1, "warmStart")) warmStart = truth(arg2);
    }
    std::fclose(config);
    if (subSize <= 0 || width % subSize != 0 || height % subSize != 0 || depth % subSize != 0) {
        std::printf("Error: Grid[%d x %d x %d] cannot be divided by subSize[%d]\n", width, height, depth, subSize);
        return false;
    }
    coarseWidth = width / subSize;
    coarseHeight = height / subSize;
    coarseDepth = depth / subSize;
    voxelSize = (float)(1.0 / width);
    if (sources.empty() && sphereSources.empty()) {
        std::printf("ERROR: no smoke source exist\n");
        return false;
    }
    if (display) {
        std::printf("resolution:\t\t%d x %d x %d\n", width, height, depth);
        std::printf("subSize:\t\t%d\n", subSize);
        std::printf("coarse:\t\t%d x %d x %d\n", coarseWidth, coarseHeight, coarseDepth);
        std::printf("Tamb:\t\t%f\n", Tamb);
        std::printf("confineCoeff:\t\t%f\n", confineCoeff);
        std::printf("velAttenCoeff:\t\t%f\n", velAttenCoeff);
        std::printf("tempAttenCoeff:\t\t%f\n", tempAttenCoeff);
        std::printf("densityAttenCoeff:\t\t%f\n", densityAttenCoeff);
        for (size_t cc = 0; cc < sources.size(); cc++)
            std::printf("Source[%d]: center(%.3f,%.3f,%.3f), size(%.3f,%.3f,%.3f), density(%.3f), temperature(%.3f)\n",
                        (int)cc, sources[cc].cx, sources[cc].cy, sources[cc].cz, sources[cc].half_xlen,
                        sources[cc].half_ylen, sources[cc].half_zlen, sources[cc].reinjectDensity,
                        sources[cc].reinjectTemperature);
        std::printf("voxelSize:\t\t%f\n", voxelSize);
        std::printf("steps:\t\t%d\n", steps);
        std::printf("stepTime:\t\t%f\n", stepTime);
        std::printf("frameCount:\t\t%d\n", frameCount);
    }
    return true;
}

bool ZQ_SmokeSimulationPara3D::IsClosedSolver() const {
    switch (globalGridSolver) {
    case CLOSED_POISSON: case CLOSED_FLUX: case CLOSED_OCTREE_POISSON: case CLOSED_OCTREE_FLUX:
    case CLOSED_OCTREE_POISSON_SOR: case TOTAL_CLOSED_POISSON_SOR: case TOTAL_CLOSED_FLUX_SOR:
    case COARSE_CLOSED_POISSON_SOR: case COARSE_CLOSED_FLUX_SOR:
        return true;
    default:
        return false;
    }
}

int ZQ_SmokeSimulationPara3D::OpenFaceMask() const {
    if (openFaces >= 0) return openFaces;
    return IsClosedSolver() ? 0 : ZQ3D_OPEN_ALL;
}

int ZQ_SmokeSimulationPara3D::WallFaceMask() const {
    if (wallFaces >= 0) return wallFaces;
    return IsClosedSolver() ? ZQ3D_OPEN_ALL : 0;
}

namespace {

void stamp_shell(bool* occ, int W, int H, int D, int shell) {
    for (int k = 0; k < D; k++)
        for (int j = 0; j < H; j++)
            for (int i = 0; i < W; i++) {
                const bool s = ((shell & 1) && i == 0) || ((shell & 2) && i == W - 1) || ((shell & 4) && j == 0) ||
                               ((shell & 8) && j == H - 1) || ((shell & 16) && k == 0) ||
                               ((shell & 32) && k == D - 1);
                if (s) occ[((size_t)k * H + j) * W + i] = true;
            }
}

}  // namespace

// ZQ_InitFineGrid3D (ZQ_SmokeSimulationUtils3D.cpp:1523-1664): empty occupancy, the closed
// solvers' one-cell wall shell (wallFaces selects a subset), the internal object, zero velocity.
bool ZQ_InitFineGrid3D(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para) {
    if (!fineGrid) return false;
    const int width = para.width, height = para.height, depth = para.depth;
    bool*& occ = fineGrid->GetOccupyPtr();
    std::memset(occ, 0, sizeof(bool) * width * height * depth);
    std::memset(fineGrid->GetPressurePtr(), 0, sizeof(float) * width * height * depth);
    stamp_shell(occ, width, height, depth, para.WallFaceMask());
    for (int k = 0; k < depth; k++)
        for (int j = 0; j < height; j++)
            for (int i = 0; i < width; i++) {
                bool s = false;
                switch (para.intobj) {
                case ZQ_SmokeSimulationPara3D::INT_OBJ_SPHERE:
                    s = (k + 0.5 - depth / 2) * (k + 0.5 - depth / 2) + (j + 0.5 - height / 2) * (j + 0.5 - height / 2) +
                            (i + 0.5 - width / 2) * (i + 0.5 - width / 2) <
                        width * width / 400.0;
                    break;
                case ZQ_SmokeSimulationPara3D::INT_OBJ_BOX:
                    s = std::fabs(k + 0.5 - depth / 2) < depth / 20 && std::fabs(j + 0.5 - height / 2) < height / 20 &&
                        std::fabs(i + 0.5 - width / 2) < width / 10;
                    break;
                case ZQ_SmokeSimulationPara3D::INT_OBJ_SPECIAL1:
                    s = (j == height / 2 + 13 || j == height / 2 + 14) && (i % 4 == 0 || i % 4 == 3) && i >= 4 &&
                        i < width - 4;
                    break;
                default:
                    break;
                }
                if (s) occ[((size_t)k * height + j) * width + i] = true;
            }
    if (para.intobj == ZQ_SmokeSimulationPara3D::INT_OBJ_OVAL) {
        // the reference runs i over height and j over width and writes occupy[k][j][i]
        // (ZQ_SmokeSimulationUtils3D.cpp:1618-1636); writes outside the grid are dropped here
        const size_t N = (size_t)width * height * depth;
        for (int k = 0; k < depth; k++)
            for (int i = 0; i < height; i++)
                for (int j = 0; j < width; j++) {
                    const double a = std::sqrt((i + 0.5 - 3.0 * width / 7) * (i + 0.5 - 3.0 * width / 7) +
                                               (j + 0.5 - height / 2) * (j + 0.5 - height / 2) +
                                               (k + 0.5 - depth / 2) * (k + 0.5 - depth / 2));
                    const double b = std::sqrt((i + 0.5 - 4.0 * width / 7) * (i + 0.5 - 4.0 * width / 7) +
                                               (j + 0.5 - height / 2) * (j + 0.5 - height / 2) +
                                               (k + 0.5 - depth / 2) * (k + 0.5 - depth / 2));
                    const size_t idx = (size_t)k * height * width + (size_t)j * width + i;
                    if (a + b < width / 5.0 && idx < N) occ[idx] = true;
                }
    }
    std::memset(fineGrid->GetVelocityUptr(), 0, sizeof(float) * (width + 1) * height * depth);
    std::memset(fineGrid->GetVelocityVptr(), 0, sizeof(float) * width * (height + 1) * depth);
    std::memset(fineGrid->GetVelocityWptr(), 0, sizeof(float) * width * height * (depth + 1));
    return true;
}

// ZQ_ApplyMovingObject3D (ZQ_SmokeSimulationUtils3D.cpp:1933-2046): clear occupancy, stamp each
// object's mask at pos - centre, set the six faces of each solid voxel to the object's velocity
// (voxels per step times h/dt, truncated to an integer as in the reference), re-add the shell.
bool ZQ_ApplyMovingObject3D(ZQ_SmokeGrid3D* fineGrid, const std::vector<ZQ_SmokeMovingObject3D*> mvobjs,
                            const ZQ_SmokeSimulationPara3D& para) {
    if (!fineGrid) return false;
    const int width = para.width, height = para.height, depth = para.depth;
    bool*& occ = fineGrid->GetOccupyPtr();
    float*& u = fineGrid->GetVelocityUptr();
    float*& v = fineGrid->GetVelocityVptr();
    float*& w = fineGrid->GetVelocityWptr();
    std::memset(occ, 0, sizeof(bool) * width * height * depth);
    for (size_t n = 0; n < mvobjs.size(); n++) {
        const ZQ_SmokeMovingObject3D* obj = mvobjs[n];
        if (!obj || !obj->GetOccupyPtr()) continue;
        int cx, cy, cz, posx, posy, posz, sx, sy, sz, cu, cv, cw;
        obj->GetCenter(cx, cy, cz);
        obj->GetPos(posx, posy, posz);
        obj->GetSize(sx, sy, sz);
        obj->GetCurVel(cu, cv, cw);
        const double scale = para.voxelSize / para.stepTime;
        cu = (int)(cu * scale);
        cv = (int)(cv * scale);
        cw = (int)(cw * scale);
        const bool* mask = obj->GetOccupyPtr();
        for (int sk = 0; sk < sz; sk++)
            for (int sj = 0; sj < sy; sj++)
                for (int si = 0; si < sx; si++) {
                    if (!mask[((size_t)sk * sy + sj) * sx + si]) continue;
                    const int ix = si + posx - cx, iy = sj + posy - cy, iz = sk + posz - cz;
                    if (ix < 0 || ix >= width || iy < 0 || iy >= height || iz < 0 || iz >= depth) continue;
                    occ[((size_t)iz * height + iy) * width + ix] = true;
                    u[((size_t)iz * height + iy) * (width + 1) + ix] = (float)cu;
                    u[((size_t)iz * height + iy) * (width + 1) + ix + 1] = (float)cu;
                    v[((size_t)iz * (height + 1) + iy) * width + ix] = (float)cv;
                    v[((size_t)iz * (height + 1) + iy + 1) * width + ix] = (float)cv;
                    w[((size_t)iz * height + iy) * width + ix] = (float)cw;
                    w[((size_t)(iz + 1) * height + iy) * width + ix] = (float)cw;
                }
    }
    stamp_shell(occ, width, height, depth, para.WallFaceMask());
    return true;
}

// ZQ_AddForce3D (ZQ_SmokeSimulationUtils3D.cpp:2049-2092) -> AddForce3D; with a nonzero
// gravityCoeff the extended entry point adds the density term.
bool ZQ_AddForce3D(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para) {
    if (!fineGrid) return false;
    const int W = fineGrid->GetWidth(), H = fineGrid->GetHeight(), D = fineGrid->GetDepth();
    float ms;
    if (para.gravityCoeff != 0.0f) {
        ms = ZQ3D_AddForce3D_ex(fineGrid->GetVelocityUptr(), fineGrid->GetVelocityVptr(), fineGrid->GetVelocityWptr(),
                                fineGrid->GetOccupyPtr(), fineGrid->GetTemperaturePtr(), fineGrid->GetDensityPtr(),
                                nullptr, para.voxelSize, para.stepTime, para.buoyancyCoeff, para.gravityCoeff,
                                para.confineCoeff, para.Tamb, W, H, D);
    } else {
        const ZQ_CUDA_AddForce3D::AddForceType type = para.addforce_type == ZQ_SmokeSimulationPara3D::ADD_FORCE_SLICE
                                                          ? ZQ_CUDA_AddForce3D::ADD_FORCE_SLICE
                                                          : ZQ_CUDA_AddForce3D::ADD_FORCE_ENTIRE;
        ms = ZQ_CUDA_AddForce3D::AddForce3D(fineGrid->GetVelocityUptr(), fineGrid->GetVelocityVptr(),
                                            fineGrid->GetVelocityWptr(), fineGrid->GetOccupyPtr(),
                                            fineGrid->GetTemperaturePtr(), para.voxelSize, para.stepTime,
                                            para.buoyancyCoeff, para.confineCoeff, para.Tamb, W, H, D, type);
    }
    if (ms < 0) {
        std::printf("add force failed: %s\n", zq3d_last_error());
        return false;
    }
    std::printf("add force gpu cost = %f\n", 0.001 * ms);
    return true;
}

// ZQ_Attenuation3D (ZQ_SmokeSimulationUtils3D.cpp:2094-2117)
bool ZQ_Attenuation3D(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para) {
    if (!fineGrid) return false;
    const float attenVel = std::exp(-para.stepTime * para.velAttenCoeff);
    const float attenDen = std::exp(-para.stepTime * para.densityAttenCoeff);
    const float attenTemp = std::exp(-para.stepTime * para.tempAttenCoeff);
    const float ms = ZQ_CUDA_Attenuation3D::Attenuation3D(
        fineGrid->GetVelocityUptr(), fineGrid->GetVelocityVptr(), fineGrid->GetVelocityWptr(),
        fineGrid->GetTemperaturePtr(), fineGrid->GetDensityPtr(), fineGrid->GetOccupyPtr(), attenVel, attenTemp,
        attenDen, para.width, para.height, para.depth);
    if (ms < 0) {
        std::printf("attenuation failed: %s\n", zq3d_last_error());
        return false;
    }
    std::printf("attenuation gpu cost = %f\n", 0.001 * ms);
    return true;
}

bool map_coarse_solver(const ZQ_SmokeSimulationPara3D& para, zq3d_config* c) {
    const GlobalSolverType3D t = para.globalGridSolver;
    const bool flux = t == COARSE_OPEN_FLUX_SOR || t == COARSE_CLOSED_FLUX_SOR || t == TOTAL_OPEN_FLUX_SOR ||
                      t == TOTAL_CLOSED_FLUX_SOR;
    if (!flux && t != COARSE_OPEN_POISSON_SOR && t != COARSE_CLOSED_POISSON_SOR) return false;
    const bool closed = t == COARSE_CLOSED_FLUX_SOR || t == TOTAL_CLOSED_FLUX_SOR;
    c->sub_size = (t == TOTAL_OPEN_FLUX_SOR || t == TOTAL_CLOSED_FLUX_SOR) ? 1 : para.subSize;
    c->coarse_flux = flux ? 1 : 0;
    c->sor.open_mask = closed ? 0 : ZQ3D_OPEN_ALL;
    c->sor.closed_mean = closed ? 1 : 0;
    c->sor.mixed_precision = 0;
    c->sor.iters = flux ? ((para.maxIter - 1) / para.innerIter + 1) * para.innerIter : para.maxIter;
    return true;
}

bool map_octree_solver(const ZQ_SmokeSimulationPara3D& para, zq3d_config* c) {
    const GlobalSolverType3D t = para.globalGridSolver;
    if (t != OPEN_OCTREE_POISSON_SOR && t != CLOSED_OCTREE_POISSON_SOR) return false;
    c->octree = 1;
    c->octree_thresh[0] = para.octree_thresh1;
    c->octree_thresh[1] = para.octree_thresh2;
    c->octree_thresh[2] = para.octree_thresh3;
    c->sub_size = 0;
    c->coarse_flux = 0;
    c->sor.open_mask = ZQ3D_OPEN_ALL;
    c->sor.closed_mean = t == CLOSED_OCTREE_POISSON_SOR ? 1 : 0;
    c->sor.mixed_precision = 0;
    c->sor.iters = para.maxIter;
    return true;
}

// ZQ_SmokeSimulationPara3D order -> ZQ_CUDA_Advection3D order (ZQ_SmokeSimulationUtils3D.cpp:2125-2148, 3802-3819)
bool map_vel_type(ZQ_SmokeSimulationPara3D::AdvectVelocityType t, ZQ_CUDA_Advection3D::AdvectVelocityType* out) {
    switch (t) {
    case ZQ_SmokeSimulationPara3D::ADV_VEL_INMAC_OUTMAC: *out = ZQ_CUDA_Advection3D::ADV_VEL_INMAC_OUTMAC; return true;
    case ZQ_SmokeSimulationPara3D::ADV_VEL_INMAC_OUTMAC_BFECC: *out = ZQ_CUDA_Advection3D::ADV_VEL_INMAC_OUTMAC_BFECC; return true;
    case ZQ_SmokeSimulationPara3D::ADV_VEL_INREG_OUTREG: *out = ZQ_CUDA_Advection3D::ADV_VEL_INREG_OUTREG; return true;
    case ZQ_SmokeSimulationPara3D::ADV_VEL_INREG_OUTREG_BFECC: *out = ZQ_CUDA_Advection3D::ADV_VEL_INREG_OUTREG_BFECC; return true;
    case ZQ_SmokeSimulationPara3D::ADV_VEL_INREG_OUTMAC: *out = ZQ_CUDA_Advection3D::ADV_VEL_INREG_OUTMAC; return true;
    case ZQ_SmokeSimulationPara3D::ADV_VEL_INREG_OUTMAC_BFECC: *out = ZQ_CUDA_Advection3D::ADV_VEL_INREG_OUTMAC_BFECC; return true;
    }
    return false;
}
namespace {
bool map_sca_type(ZQ_SmokeSimulationPara3D::AdvectScalarType t, ZQ_CUDA_Advection3D::AdvectScalarType* out) {
    switch (t) {
    case ZQ_SmokeSimulationPara3D::ADV_SCA_MAC: *out = ZQ_CUDA_Advection3D::ADV_SCA_MAC; return true;
    case ZQ_SmokeSimulationPara3D::ADV_SCA_MAC_BFECC: *out = ZQ_CUDA_Advection3D::ADV_SCA_MAC_BFECC; return true;
    case ZQ_SmokeSimulationPara3D::ADV_SCA_REG: *out = ZQ_CUDA_Advection3D::ADV_SCA_REG; return true;
    case ZQ_SmokeSimulationPara3D::ADV_SCA_REG_BFECC: *out = ZQ_CUDA_Advection3D::ADV_SCA_REG_BFECC; return true;
    }
    return false;
}
}  // namespace

// ZQ_AdvectVelocity3D (ZQ_SmokeSimulationUtils3D.cpp:2119-2168)
bool ZQ_AdvectVelocity3D(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para) {
    if (!fineGrid) return false;
    ZQ_CUDA_Advection3D::AdvectVelocityType type;
    if (!map_vel_type(para.advVelType, &type)) return false;
    const float deltatt = para.stepTime / para.steps;
    ZQ_CUDA_Advection3D::ZQ_Cuda_Prepare_Advection(para.width, para.height, para.depth, para.voxelSize, para.steps,
                                                   deltatt);
    const float ms = ZQ_CUDA_Advection3D::Advect_Velocity(fineGrid->GetVelocityUptr(), fineGrid->GetVelocityVptr(),
                                                          fineGrid->GetVelocityWptr(), fineGrid->GetOccupyPtr(), true,
                                                          type);
    if (ms < 0) {
        std::printf("velocity advection failed: %s\n", zq3d_last_error());
        return false;
    }
    std::printf("velocity advection gpu cost = %f\n", 0.001 * ms);
    return true;
}

namespace {
// extended solve when a setting the reference's call cannot express is in use
bool needs_ex(const ZQ_SmokeSimulationPara3D& para, int default_open) {
    return para.omega != 1.0f || para.mixedPrecision || para.OpenFaceMask() != default_open;
}
zq3d_sor_opts sor_opts(const ZQ_SmokeSimulationPara3D& para, int closed_mean) {
    zq3d_sor_opts o{};
    o.open_mask = para.OpenFaceMask();
    o.closed_mean = closed_mean;
    o.omega = para.omega;
    o.iters = para.maxIter;
    o.mixed_precision = para.mixedPrecision ? 1 : 0;
    o.max_outer = para.maxOuter;
    o.tol = para.sorTol;
    return o;
}
}  // namespace

// ZQ_SolveOpenPressure3DSOR (ZQ_SmokeSimulationUtils3D.cpp:2349-2366)
bool ZQ_SolveOpenPressure3DSOR(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para) {
    if (!fineGrid) return false;
    float ms;
    if (needs_ex(para, ZQ3D_OPEN_ALL)) {
        const zq3d_sor_opts o = sor_opts(para, 0);
        zq3d_proj_stats st{};
        ms = ZQ3D_SolvePoissonRedBlack_ex(fineGrid->GetVelocityUptr(), fineGrid->GetVelocityVptr(),
                                          fineGrid->GetVelocityWptr(), fineGrid->GetOccupyPtr(), fineGrid->GetWidth(),
                                          fineGrid->GetHeight(), fineGrid->GetDepth(), &o, &st);
        if (ms >= 0) std::printf("projection: %d sweeps, max|div| %g\n", st.sweeps, st.max_div);
    } else {
        ms = ZQ_CUDA_PoissonSolver3D::SolveOpenPoissonRedBlackwithOccupy3D_MAC(
            fineGrid->GetVelocityUptr(), fineGrid->GetVelocityVptr(), fineGrid->GetVelocityWptr(),
            fineGrid->GetOccupyPtr(), fineGrid->GetWidth(), fineGrid->GetHeight(), fineGrid->GetDepth(), para.maxIter);
    }
    if (ms < 0) {
        std::printf("open pressure solve failed: %s\n", zq3d_last_error());
        return false;
    }
    return true;
}

// ZQ_SolveClosedPressure3DSOR (ZQ_SmokeSimulationUtils3D.cpp:2368-2440): the mean divergence over
// fluid cells is accumulated on the host in FP32 in cell order, then the closed solve runs.
bool ZQ_SolveClosedPressure3DSOR(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para) {
    if (!fineGrid) return false;
    const int width = para.width, height = para.height, depth = para.depth;
    const bool* occupy = fineGrid->GetOccupyPtr();
    const float* U = fineGrid->GetVelocityUptr();
    const float* V = fineGrid->GetVelocityVptr();
    const float* Wf = fineGrid->GetVelocityWptr();
    float div_per_volume = 0.0f;
    int count = 0;
    for (int k = 0; k < depth; k++)
        for (int j = 0; j < height; j++)
            for (int i = 0; i < width; i++) {
                if (occupy[((size_t)k * height + j) * width + i]) continue;
                count++;
                div_per_volume += U[((size_t)k * height + j) * (width + 1) + i + 1] -
                                  U[((size_t)k * height + j) * (width + 1) + i] +
                                  V[((size_t)k * (height + 1) + j + 1) * width + i] -
                                  V[((size_t)k * (height + 1) + j) * width + i] +
                                  Wf[((size_t)(k + 1) * height + j) * width + i] - Wf[((size_t)k * height + j) * width + i];
            }
    if (count == 0) return false;
    div_per_volume /= count;
    float ms;
    if (needs_ex(para, 0)) {
        // the extended solve subtracts the FP64 mean it computes itself
        const zq3d_sor_opts o = sor_opts(para, 1);
        zq3d_proj_stats st{};
        ms = ZQ3D_SolvePoissonRedBlack_ex(fineGrid->GetVelocityUptr(), fineGrid->GetVelocityVptr(),
                                          fineGrid->GetVelocityWptr(), fineGrid->GetOccupyPtr(), width, height, depth, &o,
                                          &st);
        if (ms >= 0) std::printf("projection: %d sweeps, max|div| %g\n", st.sweeps, st.max_div);
    } else {
        ms = ZQ_CUDA_PoissonSolver3D::SolveClosedPoissonRedBlackwithOccupy3D_MAC(
            fineGrid->GetVelocityUptr(), fineGrid->GetVelocityVptr(), fineGrid->GetVelocityWptr(),
            fineGrid->GetOccupyPtr(), div_per_volume, width, height, depth, para.maxIter);
    }
    if (ms < 0) {
        std::printf("closed pressure solve failed: %s\n", zq3d_last_error());
        return false;
    }
    return true;
}

// ZQ_AdvectScalar3D (ZQ_SmokeSimulationUtils3D.cpp:3796-3839): relies on the grid constants set
// by the preceding ZQ_AdvectVelocity3D, and passes the same arrays as input and output.
bool ZQ_AdvectScalar3D(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para) {
    if (!fineGrid) return false;
    ZQ_CUDA_Advection3D::AdvectScalarType type;
    if (!map_sca_type(para.advScaType, &type)) return false;
    float*& T = fineGrid->GetTemperaturePtr();
    float*& rho = fineGrid->GetDensityPtr();
    const float ms = ZQ_CUDA_Advection3D::Advect_Scalar(fineGrid->GetVelocityUptr(), fineGrid->GetVelocityVptr(),
                                                        fineGrid->GetVelocityWptr(), fineGrid->GetOccupyPtr(), T, rho, T,
                                                        rho, type);
    if (ms < 0) {
        std::printf("scalar advection failed: %s\n", zq3d_last_error());
        return false;
    }
    std::printf("scalar advection gpu cost = %f\n", 0.001 * ms);
    return true;
}

// ZQ_ReinjectTemperatureDensity3D (ZQ_SmokeSimulationUtils3D.cpp:3842-3882), host loop; sphere
// sources after the box sources.
bool ZQ_ReinjectTemperatureDensity3D(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para) {
    if (!fineGrid) return false;
    const int width = para.width, height = para.height, depth = para.depth;
    float*& density = fineGrid->GetDensityPtr();
    float*& temperature = fineGrid->GetTemperaturePtr();
    for (size_t cc = 0; cc < para.sources.size(); cc++) {
        const ZQ_SmokeSource3D& s = para.sources[cc];
        const float cx = s.cx * width, cy = s.cy * height, cz = s.cz * depth;
        const float hx = s.half_xlen * width, hy = s.half_ylen * height, hz = s.half_zlen * depth;
        for (int k = 0; k < depth; k++)
            for (int j = 0; j < height; j++)
                for (int i = 0; i < width; i++)
                    if (std::fabs(i + 0.5f - cx) <= hx && std::fabs(j + 0.5f - cy) <= hy &&
                        std::fabs(k + 0.5f - cz) <= hz) {
                        const size_t o = ((size_t)k * height + j) * width + i;
                        density[o] = s.reinjectDensity * (para.randDensity ? std::rand() % 256 / 255.0f : 1.0f);
                        temperature[o] = s.reinjectTemperature;
                    }
    }
    for (size_t cc = 0; cc < para.sphereSources.size(); cc++) {
        const ZQ_SmokeSphereSource3D& s = para.sphereSources[cc];
        const float x0 = s.cx * width, y0 = s.cy * height, z0 = s.cz * depth, rr = s.radius * width;
        const float r2 = rr * rr;
        for (int k = 0; k < depth; k++)
            for (int j = 0; j < height; j++)
                for (int i = 0; i < width; i++) {
                    const float dx = i + 0.5f - x0, dy = j + 0.5f - y0, dz = k + 0.5f - z0;
                    if (dx * dx + dy * dy + dz * dz <= r2) {
                        const size_t o = ((size_t)k * height + j) * width + i;
                        density[o] = s.reinjectDensity;
                        temperature[o] = s.reinjectTemperature;
                    }
                }
    }
    return true;
}

void MACtoRegular4(int width, int height, int depth, const float* mac_u, const float* mac_v, const float* mac_w,
                   float* vel4) {
    ZQ3D_MACtoRegular4(mac_u, mac_v, mac_w, vel4, width, height, depth);
}
