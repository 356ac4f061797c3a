// World: the reference's domain class (ch3/ver2/World.h:38-147, ch2/World.h) over the GPU
// library.  The node fields phi, rho, node_vol, ef and object_id live on the device; the
// public Field members of the same names are host mirrors refreshed by updateHost()
// (the updateHostPhi() boundary of ch9/CUDA/PotentialSolver.cu:323-329).
#ifndef PICAMD_HOST_WORLD_H
#define PICAMD_HOST_WORLD_H

#include <chrono>
#include <memory>
#include <random>
#include <vector>

#include "Field.h"
#include "picamd.h"

class Species;

namespace Const {
const double EPS_0 = 8.85418782e-12;  // C/(V*m), vacuum permittivity
const double QE = 1.602176565e-19;    // C, electron charge
const double AMU = 1.660538921e-27;   // kg, atomic mass unit
const double ME = 9.10938215e-31;     // kg, electron mass
const double K = 1.380648e-23;        // J/K, Boltzmann constant
const double PI = 3.141592653;        // pi
const double EvToK = QE / K;          // 1eV in K ~ 11604
}  // namespace Const

// Uniform [0,1) generator, std::mt19937 + uniform_real_distribution<double> as the
// reference's Rnd (World.h:24-33).  The reference seeds from std::random_device; seed()
// makes a run reproducible (host loads then match the reference draw for draw).
class Rnd {
public:
    Rnd() : mt_gen{std::random_device()()}, rnd_dist{0, 1.0} {}
    double operator()() { return rnd_dist(mt_gen); }
    void seed(unsigned s) {
        mt_gen.seed(s);
        rnd_dist.reset();
    }
    unsigned long long next64() { return ((unsigned long long)mt_gen() << 32) | mt_gen(); }

protected:
    std::mt19937 mt_gen;
    std::uniform_real_distribution<double> rnd_dist;
};

extern Rnd rnd;

// Throws on a library error (the reference has no error paths on these calls).
void pic_check(int rc, const char *what);

class World {
public:
    World(int ni, int nj, int nk);
    ~World();
    World(const World &) = delete;
    World &operator=(const World &) = delete;

    void setExtents(const double3 x0, const double3 xm);

    double3 getX0() const { return double3(x0); }
    double3 getXm() const { return double3(xm); }
    double3 getXc() const { return double3(xc); }
    double3 getDh() const { return double3(dh); }

    int getTs() const { return ts; }
    double getTime() const { return time; }
    double getWallTime();
    double getDt() const { return dt; }
    bool isLastTimeStep() const { return ts == num_ts - 1; }

    bool inBounds(double3 pos) {
        for (int i = 0; i < 3; i++)
            if (pos[i] < x0[i] || pos[i] >= xm[i]) return false;
        return true;
    }

    void setTime(double dt, int num_ts) {
        this->dt = dt;
        this->num_ts = num_ts;
    }
    // pre-increment: the loop body runs num_ts+1 times (World.h:69)
    bool advanceTime() {
        time += dt;
        ts++;
        return ts <= num_ts;
    }

    bool steadyState(std::vector<Species> &species);

    double3 XtoL(double3 x) const {
        double3 lc;
        for (int c = 0; c < 3; c++) lc[c] = (x[c] - x0(c)) / dh(c);
        return lc;
    }
    double3 pos(double3 lc) { return x0 + dh * lc; }
    double3 pos(int i, int j, int k) {
        double3 x{(double)i, (double)j, (double)k};
        return pos(x);
    }
    int U(int i, int j, int k) { return object_id.U(i, j, k); }

    void computeChargeDensity(std::vector<Species> &species);
    double getPE();
    void addSphere(double3 x0, double radius, double phi_sphere);
    bool inSphere(double3 x);
    void addInlet();

    // ---- GPU boundary ---------------------------------------------------------------
    void updateHost();    // device -> phi, rho, node_vol, ef, object_id mirrors
    void updateDevice();  // phi, rho mirrors -> device
    pic_world *dev() { return handle; }
    // particle arithmetic of species created from now on (PIC_MIXED or PIC_FP64)
    void setPrecision(int p) { precision = p; }
    int getPrecision() const { return precision; }
    // absorbing walls once a sphere or inlet exists (the ch3 geometry), else reflecting
    int particleBoundary() const { return absorbing ? PIC_BC_ABSORB : PIC_BC_REFLECT; }
    void setParticleBoundary(int bc) {
        absorbing = (bc == PIC_BC_ABSORB);
        bc_set = true;
    }
    unsigned long long sourceSeed() { return rnd.next64(); }

    const int ni, nj, nk;
    const int3 nn;

    Field phi;
    Field rho;
    Field node_vol;
    Field3 ef;
    FieldI object_id;

protected:
    double3 x0, dh, xm, xc;
    double dt = 0;
    double time = 0;
    int ts = -1;
    int num_ts = 0;
    double3 sphere_x0{0, 0, 0};
    double sphere_rad2 = 0;
    std::chrono::time_point<std::chrono::high_resolution_clock> time_start;
    // steadyState(): the totals it compared last (real count, |z momentum|, PE + KE)
    struct Totals {
        double count = 0, pz = 0, energy = 0;
    };
    bool steady_state = false;
    Totals previous;

    pic_world *handle = nullptr;
    int precision = PIC_MIXED;
    bool absorbing = false;
    bool bc_set = false;
};

#endif
