// Device injection behind the reference's source classes (see Source.h).
#include "Source.h"

static void drift_frame(Species &sp, double v_drift) {
    const double u[3] = {0.0, 0.0, v_drift};
    pic_check(pic_species_set_frame_velocity(sp.dev(), u), "Source: frame velocity");
}

ColdBeamSource::ColdBeamSource(Species &species, World &world, double v_drift, double den)
    : sp{&species}, world{&world}, v_drift{v_drift}, den{den}, seed{world.sourceSeed()} {
    drift_frame(species, v_drift);
}

WarmBeamSource::WarmBeamSource(Species &species, World &world, double v_drift, double den, double T)
    : sp{&species}, world{&world}, v_drift{v_drift}, den{den}, T{T}, seed{world.sourceSeed()} {
    drift_frame(species, v_drift);
}

void ColdBeamSource::sample() {
    long n = 0;
    pic_check(pic_species_inject_beam(sp->dev(), v_drift, den, 0.0, world->getDt(), seed, counter++, &n),
              "ColdBeamSource::sample");
}

void WarmBeamSource::sample() {
    long n = 0;
    pic_check(pic_species_inject_beam(sp->dev(), v_drift, den, T, world->getDt(), seed, counter++, &n),
              "WarmBeamSource::sample");
}
