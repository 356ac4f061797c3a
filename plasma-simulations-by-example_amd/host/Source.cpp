// Device injection behind the reference's source classes (see Source.h).
#include "Source.h"

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
