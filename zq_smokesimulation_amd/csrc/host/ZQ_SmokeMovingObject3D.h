// ZQ_SmokeMovingObject3D -- voxel-mask moving obstacle with the reference's interface and file
// format (ZQ_SmokeMovingObject3D.h:6-37, .cpp:74-124).
#pragma once

#include <vector>

class ZQ_SmokeMovingObject3D {
public:
    class Velocity {
    public:
        int x, y, z;
        Velocity(int u = 0, int v = 0, int w = 0) : x(u), y(v), z(w) {}
    };

    ZQ_SmokeMovingObject3D();
    ~ZQ_SmokeMovingObject3D();
    ZQ_SmokeMovingObject3D(const ZQ_SmokeMovingObject3D&) = delete;
    ZQ_SmokeMovingObject3D& operator=(const ZQ_SmokeMovingObject3D&) = delete;

    int cx, cy, cz;        // centre of the mask (voxels)
    int posx, posy, posz;  // position in the grid (voxels)
    int sizex, sizey, sizez;
    int curstep;
    std::vector<Velocity> vels;  // voxels per step
    bool* occupy;

    void GetCenter(int& cx, int& cy, int& cz) const;
    void GetPos(int& posx, int& posy, int& posz) const;
    void GetSize(int& sizex, int& sizey, int& sizez) const;
    void GetCurVel(int& u, int& v, int& w) const;
    const bool* GetOccupyPtr() const { return occupy; }
    void UpdateOneStep();
};

// file: count, then per object: size xyz, centre xyz, position xyz, step count, one velocity
// triple per step, size.x*size.y*size.z mask values.  Appends to mvobjs; false if unreadable.
bool LoadMovingObjects3D(const char* file, std::vector<ZQ_SmokeMovingObject3D*>& mvobjs);
