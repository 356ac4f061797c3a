// Species over libpicamd (see Species.h).
#include "Species.h"

#include <algorithm>
#include <map>

struct Species::Impl {
    pic_species *h = nullptr;
    double dt_hint = 0;
    std::vector<double> qpos, qvel, qmpw;  // queued addParticle calls
    ~Impl() { pic_species_destroy(h); }
};

namespace {
// species of each world, for flushAll before E changes
std::map<World *, std::vector<std::weak_ptr<Species::Impl>>> &registry() {
    static std::map<World *, std::vector<std::weak_ptr<Species::Impl>>> r;
    return r;
}

void flush_impl(Species::Impl &m, World &w) {
    if (m.qmpw.empty()) return;
    long added = 0;
    pic_check(pic_species_add_particles(m.h, (long)m.qmpw.size(), m.qpos.data(), m.qvel.data(), m.qmpw.data(),
                                        w.getDt(), &added),
              "pic_species_add_particles");
    m.qpos.clear();
    m.qvel.clear();
    m.qmpw.clear();
}
}  // namespace

Species::Species(std::string name, double mass, double charge, double mpw0, World &world)
    : name(name), mass(mass), charge(charge), mpw0(mpw0), den(world.ni, world.nj, world.nk),
      den_ave(world.ni, world.nj, world.nk), world(&world), impl(std::make_shared<Impl>()) {
    pic_check(pic_species_create(world.dev(), mass, charge, mpw0, world.getPrecision(), 1 << 16, 0, &impl->h),
              "pic_species_create");
    registry()[&world].push_back(impl);
}

Species::Species(std::string name, double mass, double charge, World &world)
    : Species(name, mass, charge, 0.0, world) {}

pic_species *Species::dev() {
    flush();
    return impl->h;
}

void Species::flush() { flush_impl(*impl, *world); }

void Species::flushAll(World &w) {
    auto &v = registry()[&w];
    for (auto it = v.begin(); it != v.end();) {
        if (auto p = it->lock()) {
            flush_impl(*p, w);
            ++it;
        } else {
            it = v.erase(it);
        }
    }
}

size_t Species::getNp() {
    long n = 0;
    pic_check(pic_species_get_np(dev(), &n), "pic_species_get_np");
    return (size_t)n;
}

void Species::diag(double out[5]) { pic_check(pic_species_diag(dev(), out), "pic_species_diag"); }

double Species::getRealCount() {
    double d[5];
    diag(d);
    return d[0];
}

double3 Species::getMomentum() {
    double d[5];
    diag(d);
    return double3(d[1], d[2], d[3]);
}

double Species::getKE() {
    double d[5];
    diag(d);
    return d[4];
}

void Species::advance() {
    pic_check(pic_species_advance(dev(), world->particleBoundary(), world->getDt()), "pic_species_advance");
}

void Species::computeNumberDensity() {
    pic_check(pic_species_compute_number_density(dev()), "pic_species_compute_number_density");
}

// Species::addParticle (ch3/ver2/Species.cpp:65-81): out-of-bounds positions are dropped,
// the rewind happens on the device when the queue is flushed
void Species::addParticle(double3 pos, double3 vel, double mpw) {
    if (!world->inBounds(pos)) return;
    for (int c = 0; c < 3; c++) {
        impl->qpos.push_back(pos[c]);
        impl->qvel.push_back(vel[c]);
    }
    impl->qmpw.push_back(mpw);
}

// Species::loadParticlesBox (ch2/Species.cpp:74-97): the same three draws per particle from
// the global rnd, so a seeded run reproduces the reference's positions exactly
void Species::loadParticlesBox(double3 x1, double3 x2, double num_den, int num_mp) {
    double box_vol = (x2[0] - x1[0]) * (x2[1] - x1[1]) * (x2[2] - x1[2]);
    double num_real = num_den * box_vol;
    double mpw = num_real / num_mp;
    impl->qpos.reserve(impl->qpos.size() + 3 * (size_t)num_mp);
    for (int p = 0; p < num_mp; p++) {
        double3 pos;
        for (int c = 0; c < 3; c++) pos[c] = x1[c] + rnd() * (x2[c] - x1[c]);
        addParticle(pos, double3(0, 0, 0), mpw);
    }
}

// Species::loadParticlesBoxQS (ch2/Species.cpp:101-141): lattice of num_mp points per axis,
// half/quarter/eighth weights on faces, edges and corners
void Species::loadParticlesBoxQS(double3 x1, double3 x2, double num_den, int3 num_mp) {
    double box_vol = (x2[0] - x1[0]) * (x2[1] - x1[1]) * (x2[2] - x1[2]);
    int num_mp_tot = (num_mp[0] - 1) * (num_mp[1] - 1) * (num_mp[2] - 1);
    double mpw = num_den * box_vol / num_mp_tot;
    double d[3];
    for (int c = 0; c < 3; c++) d[c] = (x2[c] - x1[c]) / (num_mp[c] - 1);
    for (int i = 0; i < num_mp[0]; i++)
        for (int j = 0; j < num_mp[1]; j++)
            for (int k = 0; k < num_mp[2]; k++) {
                int idx[3] = {i, j, k};
                double3 pos;
                double w = 1;
                for (int c = 0; c < 3; c++) {
                    pos[c] = x1[c] + idx[c] * d[c];
                    if (pos[c] == x2[c]) pos[c] -= 1e-4 * d[c];  // keep the max faces inside
                    if (idx[c] == 0 || idx[c] == num_mp[c] - 1) w *= 0.5;
                }
                addParticle(pos, double3(0, 0, 0), mpw * w);
            }
}

void Species::updateAverages() { pic_check(pic_species_update_averages(dev()), "pic_species_update_averages"); }

void Species::syncHost(bool with_particles) {
    size_t n = (size_t)den.ni * den.nj * den.nk;
    std::vector<double> buf(n);
    pic_check(pic_species_get_den(dev(), buf.data()), "pic_species_get_den");
    den.inflate(buf.data());
    pic_check(pic_species_get_den_ave(dev(), buf.data()), "pic_species_get_den_ave");
    den_ave.inflate(buf.data());
    if (!with_particles) return;
    size_t np = getNp();
    std::vector<double> pos(3 * np), vel(3 * np), mpw(np);
    if (np) pic_check(pic_species_get_particles(impl->h, pos.data(), vel.data(), mpw.data(), nullptr), "get particles");
    particles.clear();
    particles.reserve(np);
    for (size_t p = 0; p < np; p++)
        particles.emplace_back(double3(&pos[3 * p]), double3(&vel[3 * p]), mpw[p]);
}
