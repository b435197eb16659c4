// ZQ_SmokeMovingObject3D (reference: ZQ_SmokeMovingObject3D.cpp:1-124)
#include "ZQ_SmokeMovingObject3D.h"

#include <cstdio>

ZQ_SmokeMovingObject3D::ZQ_SmokeMovingObject3D()
    : cx(0), cy(0), cz(0), posx(0), posy(0), posz(0), sizex(0), sizey(0), sizez(0), curstep(0), occupy(nullptr) {}

ZQ_SmokeMovingObject3D::~ZQ_SmokeMovingObject3D() { delete[] occupy; }

void ZQ_SmokeMovingObject3D::GetCenter(int& x, int& y, int& z) const {
    x = cx;
    y = cy;
    z = cz;
}

void ZQ_SmokeMovingObject3D::GetPos(int& x, int& y, int& z) const {
    x = posx;
    y = posy;
    z = posz;
}

void ZQ_SmokeMovingObject3D::GetSize(int& x, int& y, int& z) const {
    x = sizex;
    y = sizey;
    z = sizez;
}

// the step's velocity, 0 once the list is exhausted
void ZQ_SmokeMovingObject3D::GetCurVel(int& u, int& v, int& w) const {
    if (curstep >= (int)vels.size() || curstep < 0) {
        u = v = w = 0;
    } else {
        u = vels[curstep].x;
        v = vels[curstep].y;
        w = vels[curstep].z;
    }
}

void ZQ_SmokeMovingObject3D::UpdateOneStep() {
    if (curstep >= 0 && curstep < (int)vels.size()) {
        posx += vels[curstep].x;
        posy += vels[curstep].y;
        posz += vels[curstep].z;
    }
    curstep++;
}

bool LoadMovingObjects3D(const char* file, std::vector<ZQ_SmokeMovingObject3D*>& mvobjs) {
    FILE* in = std::fopen(file, "r");
    if (!in) return false;
    int count = 0;
    bool ok = std::fscanf(in, "%d", &count) == 1 && count >= 0;
    for (int i = 0; ok && i < count; i++) {
        int sx, sy, sz, cx, cy, cz, px, py, pz, steps;
        ok = std::fscanf(in, "%d%d%d", &sx, &sy, &sz) == 3 && std::fscanf(in, "%d%d%d", &cx, &cy, &cz) == 3 &&
             std::fscanf(in, "%d%d%d", &px, &py, &pz) == 3 && std::fscanf(in, "%d", &steps) == 1 && sx >= 0 &&
             sy >= 0 && sz >= 0 && steps >= 0;
        if (!ok) break;
        ZQ_SmokeMovingObject3D* obj = new ZQ_SmokeMovingObject3D();
        obj->sizex = sx;
        obj->sizey = sy;
        obj->sizez = sz;
        obj->cx = cx;
        obj->cy = cy;
        obj->cz = cz;
        obj->posx = px;
        obj->posy = py;
        obj->posz = pz;
        for (int st = 0; ok && st < steps; st++) {
            int u, v, w;
            ok = std::fscanf(in, "%d%d%d", &u, &v, &w) == 3;
            obj->vels.push_back(ZQ_SmokeMovingObject3D::Velocity(u, v, w));
        }
        const int n = sx * sy * sz;
        obj->occupy = new bool[n > 0 ? n : 1];
        for (int b = 0; ok && b < n; b++) {
            int val = 0;
            ok = std::fscanf(in, "%d", &val) == 1;
            obj->occupy[b] = val != 0;
        }
        mvobjs.push_back(obj);
    }
    std::fclose(in);
    return ok;
}
