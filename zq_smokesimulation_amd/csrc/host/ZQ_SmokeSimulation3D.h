// ZQ_SmokeSimulation3D -- simulation driver with the reference's interface
// (ZQ_SmokeSimulation3D.h:13-47).
#pragma once

#include <vector>

#include "ZQ_SmokeGrid3D.h"
#include "ZQ_SmokeMovingObject3D.h"
#include "ZQ_SmokeSimulationUtils3D.h"

// guide grids drive the guiding control of the coarse solver (ZQ_SmokeGuideGrid3D.h)
class ZQ_SmokeGuideGrid3D;

class ZQ_SmokeSimulation3D {
public:
    ZQ_SmokeSimulation3D();
    ~ZQ_SmokeSimulation3D();
    ZQ_SmokeSimulation3D(const ZQ_SmokeSimulation3D&) = delete;
    ZQ_SmokeSimulation3D& operator=(const ZQ_SmokeSimulation3D&) = delete;

    // parse the config, allocate and initialise the grid (ZQ_SmokeSimulation3D.cpp:60-100)
    bool Init(const char* file);
    // frameCount frames, density of frame i written to <fold>/frame<i>.di3 (ZQ_SmokeSimulation3D.cpp:102-154)
    bool Run(const char* fold);
    // one step in the reference's stage order (ZQ_SmokeSimulation3D.cpp:156-243)
    bool UpdateOneFrame(const ZQ_SmokeGuideGrid3D* guideGrid = nullptr);

    const ZQ_SmokeSimulationPara3D& GetPara() const { return para; }
    // host grid; with deviceResident it is brought up to date by SyncToHost()
    ZQ_SmokeGrid3D* GetFineGrid() { return fineGrid; }
    bool SyncToHost();
    const zq3d_proj_stats& LastProjectionStats() const { return lastStats; }

private:
    ZQ_SmokeSimulationPara3D para;
    ZQ_SmokeGrid3D* fineGrid;
    std::vector<ZQ_SmokeMovingObject3D*> mvobjs;
    zq3d_sim* dev;        // device-resident state (deviceResident and no file-based moving objects)
    zq3d_sim* proj;       // staged path with a COARSE_* solver: projection-only device simulation
    int moving_frame;     // the reference keeps this in a function-static (ZQ_SmokeSimulation3D.cpp:161)
    bool host_current;    // host grid equals the device state
    zq3d_proj_stats lastStats;

    bool is_coarse() const;                    // a coarse-grid solve (guiding applies)
    bool device_projection() const;            // the coarse-grid and octree solves
    void set_projection(zq3d_config* c) const;
    bool _initRegular();
    bool _projection();
    bool _updateResident();
    bool _updateHost();
};
