// World over libpicamd (see World.h).
#include "World.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <stdexcept>
#include <string>

#include "Species.h"

Rnd rnd;

using namespace std;

void pic_check(int rc, const char *what) {
    if (rc == PIC_OK) return;
    throw std::runtime_error(std::string(what) + ": " + pic_last_error());
}

World::World(int ni, int nj, int nk)
    : ni{ni}, nj{nj}, nk{nk}, nn{ni, nj, nk}, phi(ni, nj, nk), rho(ni, nj, nk), node_vol(ni, nj, nk),
      ef(ni, nj, nk), object_id(ni, nj, nk) {
    time_start = chrono::high_resolution_clock::now();
    pic_check(pic_world_create(ni, nj, nk, &handle), "pic_world_create");
    // an unchanged Main.cpp never calls setPrecision: PICAMD_PRECISION=fp64 selects the FP64
    // form for it (default: mixed)
    const char *p = getenv("PICAMD_PRECISION");
    if (p && strcmp(p, "fp64") == 0) precision = PIC_FP64;
}

World::~World() { pic_world_destroy(handle); }

void World::setExtents(double3 _x0, double3 _xm) {
    x0 = _x0;
    xm = _xm;
    for (int i = 0; i < 3; i++) dh[i] = (xm(i) - x0(i)) / (nn(i) - 1);
    xc = 0.5 * (x0 + xm);
    double a[3] = {x0[0], x0[1], x0[2]}, b[3] = {xm[0], xm[1], xm[2]};
    pic_check(pic_world_set_extents(handle, a, b), "pic_world_set_extents");
}

double World::getWallTime() {
    auto now = chrono::high_resolution_clock::now();
    chrono::duration<double> d = now - time_start;
    return d.count();
}

void World::computeChargeDensity(vector<Species> &species) {
    vector<pic_species *> hs;
    for (Species &sp : species) {
        sp.flush();
        hs.push_back(sp.dev());
    }
    pic_check(pic_world_compute_charge_density(handle, hs.data(), (int)hs.size()),
              "pic_world_compute_charge_density");
}

double World::getPE() {
    double pe = 0;
    pic_check(pic_world_get_pe(handle, &pe), "pic_world_get_pe");
    return pe;
}

void World::addSphere(double3 c, double radius, double phi_sphere) {
    sphere_x0 = c;
    sphere_rad2 = radius * radius;
    double a[3] = {c[0], c[1], c[2]};
    pic_check(pic_world_add_sphere(handle, a, radius, phi_sphere), "pic_world_add_sphere");
    if (!bc_set) absorbing = true;
}

void World::addInlet() {
    pic_check(pic_world_add_inlet(handle), "pic_world_add_inlet");
    if (!bc_set) absorbing = true;
}

bool World::inSphere(double3 x) {
    double3 r = x - sphere_x0;
    double m2 = (r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    return m2 <= sphere_rad2;
}

// World::steadyState (ch3/ver2/World.cpp:129-158).  The totals over the species, from the
// device reductions, are accumulated in the reference's order (energy starts from PE); the
// state turns steady once each total has moved by less than 1e-3 of itself since the last
// call, and stays steady without further reductions.
bool World::steadyState(vector<Species> &species) {
    if (steady_state) return true;
    Totals now;
    now.energy = getPE();
    for (Species &sp : species) {
        now.count += sp.getRealCount();
        now.pz += std::abs(sp.getMomentum()[2]);
        now.energy += sp.getKE();
    }
    auto settled = [](double cur, double prev) { return std::abs((cur - prev) / cur) < 1e-3; };
    steady_state = settled(now.count, previous.count) && settled(now.pz, previous.pz) &&
                   settled(now.energy, previous.energy);
    if (steady_state) cout << "Steady state reached at time step " << ts << endl;
    previous = now;
    return steady_state;
}

void World::updateHost() {
    size_t n = (size_t)ni * nj * nk;
    vector<double> buf(3 * n);
    pic_check(pic_world_get_field(handle, PIC_FIELD_PHI, buf.data()), "get phi");
    phi.inflate(buf.data());
    pic_check(pic_world_get_field(handle, PIC_FIELD_RHO, buf.data()), "get rho");
    rho.inflate(buf.data());
    pic_check(pic_world_get_field(handle, PIC_FIELD_NODE_VOL, buf.data()), "get node_vol");
    node_vol.inflate(buf.data());
    pic_check(pic_world_get_field(handle, PIC_FIELD_OBJECT_ID, buf.data()), "get object_id");
    for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++)
            for (int k = 0; k < nk; k++) object_id[i][j][k] = (int)buf[object_id.U(i, j, k)];
    pic_check(pic_world_get_field(handle, PIC_FIELD_EF, buf.data()), "get ef");
    for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++)
            for (int k = 0; k < nk; k++) {
                size_t u = (size_t)ef.U(i, j, k);
                ef[i][j][k] = double3(buf[3 * u], buf[3 * u + 1], buf[3 * u + 2]);
            }
}

void World::updateDevice() {
    size_t n = (size_t)ni * nj * nk;
    vector<double> buf(n);
    phi.deflate(buf.data());
    pic_check(pic_world_set_field(handle, PIC_FIELD_PHI, buf.data()), "set phi");
    rho.deflate(buf.data());
    pic_check(pic_world_set_field(handle, PIC_FIELD_RHO, buf.data()), "set rho");
}
