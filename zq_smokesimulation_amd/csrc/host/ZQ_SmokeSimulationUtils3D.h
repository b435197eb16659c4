// ZQ_SmokeSimulationUtils3D -- parameters, config parser and stage functions with the reference's
// interface (ZQ_SmokeSimulationUtils3D.h:18-171), calling the gfx950 C ABI (zq_smoke3d.h).
#pragma once

#include <vector>

#include "ZQ_SmokeGrid3D.h"
#include "ZQ_SmokeMovingObject3D.h"
#include "zq_smoke3d.h"

// ZQ_SmokeSimulationUtils3D.h:18-38 (same order, same values)
enum GlobalSolverType3D {
    CLOSED_POISSON,
    CLOSED_FLUX,
    OPEN_POISSON,
    OPEN_FLUX,
    COARSE_OPEN_POISSON_SOR,
    COARSE_CLOSED_POISSON_SOR,
    COARSE_OPEN_FLUX_SOR,
    COARSE_CLOSED_FLUX_SOR,
    TOTAL_OPEN_POISSON_SOR,
    TOTAL_CLOSED_POISSON_SOR,
    TOTAL_OPEN_FLUX_SOR,
    TOTAL_CLOSED_FLUX_SOR,
    CLOSED_OCTREE_POISSON,
    CLOSED_OCTREE_FLUX,
    OPEN_OCTREE_POISSON,
    OPEN_OCTREE_FLUX,
    OPEN_OCTREE_POISSON_SOR,
    CLOSED_OCTREE_POISSON_SOR
};

// box source (ZQ_SmokeSimulationUtils3D.h:41-48): centre and half extents normalized by W/H/D
class ZQ_SmokeSource3D {
public:
    float cx, cy, cz;
    float half_xlen, half_ylen, half_zlen;
    float reinjectDensity;
    float reinjectTemperature;
};

// sphere source (not in the reference): centre normalized by W/H/D, radius in units of W
class ZQ_SmokeSphereSource3D {
public:
    float cx, cy, cz, radius;
    float reinjectDensity;
    float reinjectTemperature;
};

class ZQ_SmokeSimulationPara3D {
public:
    enum InternalObject { INT_OBJ_NULL, INT_OBJ_SPHERE, INT_OBJ_BOX, INT_OBJ_OVAL, INT_OBJ_SPECIAL1 };
    enum AdvectVelocityType {
        ADV_VEL_INMAC_OUTMAC,
        ADV_VEL_INMAC_OUTMAC_BFECC,
        ADV_VEL_INREG_OUTREG,
        ADV_VEL_INREG_OUTREG_BFECC,
        ADV_VEL_INREG_OUTMAC,
        ADV_VEL_INREG_OUTMAC_BFECC
    };
    enum AdvectScalarType { ADV_SCA_MAC, ADV_SCA_MAC_BFECC, ADV_SCA_REG, ADV_SCA_REG_BFECC };
    enum ExportType { EXPORT_DI3, EXPORT_ZQWV, EXPORT_ZQCI };
    enum AddForceType { ADD_FORCE_ENTIRE, ADD_FORCE_SLICE };

    ZQ_SmokeSimulationPara3D();

    // ---- reference parameters (ZQ_SmokeSimulationUtils3D.h:82-121) ----
    int width, height, depth;
    int coarseWidth, coarseHeight, coarseDepth;
    int subSize;
    float gravityCoeff;   // density term alpha: f_y -= alpha*rho (the reference declares it unused)
    float buoyancyCoeff;  // beta
    float confineCoeff;   // vorticity confinement eps
    std::vector<ZQ_SmokeSource3D> sources;
    float velAttenCoeff;
    float Tamb;
    float voxelSize;
    float stepTime;
    float tempAttenCoeff;
    float densityAttenCoeff;
    int steps;
    int frameCount;
    int maxIter;
    int innerIter;
    GlobalSolverType3D globalGridSolver;
    int subGridSolver;
    int octree_thresh1, octree_thresh2, octree_thresh3;
    bool useMovingObject;
    char movingObjectFile[200];
    bool useGuidingControl;
    float guidingCoeff;
    bool guidingNaiveBlend;
    char guidingFold[200];
    bool exportGuide;
    bool exportOccupy;
    bool randDensity;
    float maxDensity;
    float maxTemperature;
    InternalObject intobj;
    AdvectVelocityType advVelType;
    AdvectScalarType advScaType;
    ExportType exportType;
    float export_quality;
    AddForceType addforce_type;

    // ---- extensions (keys the reference's parser does not know) ----
    std::vector<ZQ_SmokeSphereSource3D> sphereSources;  // sphereSource cx cy cz r rho T
    float omega;          // omega: SOR relaxation (1 = the reference's Gauss-Seidel form)
    bool mixedPrecision;  // mixedPrecision true: FP64 residual refinement
    float sorTol;         // sorTol: max|div| target of the mixed-precision solve
    int maxOuter;         // maxOuter: refinement passes at most
    int openFaces;        // openFaces: domain faces with p = 0 (-1: all for open solvers, none for closed)
    int wallFaces;        // wallFaces: faces with a solid shell (-1: all for closed solvers, none for open)
    float tempNoise;      // tempNoise amp [seed]: uniform +-amp added to T once at start
    unsigned noiseSeed;
    bool deviceResident;  // deviceResident: keep the fields in HBM between stages (default true)
    bool warmStart;       // warmStart: keep the pressure between frames (mixed precision)

    bool LoadFromFile(const char* file, bool display = true);

    // derived: domain faces that are open / walled for this solver
    int OpenFaceMask() const;
    int WallFaceMask() const;
    bool IsClosedSolver() const;
};

bool ZQ_InitFineGrid3D(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para);
bool ZQ_ApplyMovingObject3D(ZQ_SmokeGrid3D* fineGrid, const std::vector<ZQ_SmokeMovingObject3D*> mvobjs,
                            const ZQ_SmokeSimulationPara3D& para);
bool ZQ_AddForce3D(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para);
bool ZQ_Attenuation3D(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para);
bool ZQ_AdvectVelocity3D(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para);
// The projection settings of the solvers that run on the coarse-grid machinery: COARSE_*_POISSON_SOR
// (both the open face-fraction solve, as ZQ_TotalProjection3D calls it, :3730-3745), COARSE_*_FLUX_SOR
// and TOTAL_*_FLUX_SOR (the flux form; TOTAL_* on the fine grid itself, s = 1), the flux solvers with
// (maxIter-1)/innerIter+1 passes of innerIter iterations (:2452-2458).  Returns false (c untouched)
// for the other solvers.
bool map_coarse_solver(const ZQ_SmokeSimulationPara3D& para, zq3d_config* c);
// The octree solvers with the red-black solve, OPEN_OCTREE_POISSON_SOR and CLOSED_OCTREE_POISSON_SOR
// (ZQ_SmokeOctreeGrid3D.cpp:1300-1353, 2850-3018): zq3d_config.octree with octree_thresh1..3 and
// maxIter passes; the closed one removes the mean divergence.  Returns false (c untouched) for the
// other solvers.
bool map_octree_solver(const ZQ_SmokeSimulationPara3D& para, zq3d_config* c);
// the config's velocity advection type in ZQ_CUDA_Advection3D's numbering (false: unknown)
bool map_vel_type(ZQ_SmokeSimulationPara3D::AdvectVelocityType t, ZQ_CUDA_Advection3D::AdvectVelocityType* out);
bool ZQ_SolveOpenPressure3DSOR(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para);
bool ZQ_SolveClosedPressure3DSOR(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para);
bool ZQ_AdvectScalar3D(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para);
bool ZQ_ReinjectTemperatureDensity3D(ZQ_SmokeGrid3D* fineGrid, const ZQ_SmokeSimulationPara3D& para);
// cell-centred (u,v,w,0) from the faces (the reference's CPU twin, ZQ_SmokeSimulationUtils3D.cpp:355-374)
void MACtoRegular4(int width, int height, int depth, const float* mac_u, const float* mac_v, const float* mac_w,
                   float* vel4);
