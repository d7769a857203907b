// ref_driver.cpp -- C-ABI wrapper over the reference's own host-compilable sources.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) in the container
// that holds the reference checkout: it includes the reference's src/common.h and
// src/designs.cu straight from that checkout (nothing is copied into this repository)
// and links its src/stl.cu, exactly the file set of the reference's own `make host-test`
// recipe (Makefile:63-65).  The output goes to oracle/_ref/ and is used to pin the oracle
// (tests/golden/make_golden.py) -- never by the product.
#include "common.h"
#include "designs.cu"

extern "C" {

int ref_build_mesh(const Genome* g, float* V, int* T, int* nv, int* nt) {
    v3 VV[MAX_V]; int TT[MAX_T][3]; MeshOut mo;
    build_mesh(g, VV, TT, &mo);
    *nv = mo.nv; *nt = mo.nt;
    for (int i = 0; i < mo.nv; i++) { V[3*i] = VV[i].x; V[3*i+1] = VV[i].y; V[3*i+2] = VV[i].z; }
    for (int i = 0; i < mo.nt; i++) for (int k = 0; k < 3; k++) T[3*i+k] = TT[i][k];
    return fits_print_volume(VV, mo.nv, PRINT_VOLUME_M - PLA_WALL_M);
}

void ref_mass_props(const Genome* g, MassProps* out) {
    v3 VV[MAX_V]; int TT[MAX_T][3]; MeshOut mo;
    build_mesh(g, VV, TT, &mo);
    mass_props(g, VV, TT, mo.nt, out);
}

void ref_wing_pt(const Genome* g, float u, float xc, float zoff, float scale, float* out3) {
    v3 p = wing_pt(g, u, xc, zoff, scale);
    out3[0] = p.x; out3[1] = p.y; out3[2] = p.z;
}

void ref_scenario(int seed, ScenarioDesc* out) { *out = scenario_desc(seed); }

float ref_u01(int seed, int lane) { return scenario_u01(seed, lane); }

void ref_printed(const MassProps* src, int seed, MassProps* out, float* shift3) {
    v3 sh; *out = printed_mass_props(*src, scenario_desc(seed), &sh);
    shift3[0] = sh.x; shift3[1] = sh.y; shift3[2] = sh.z;
}

void ref_wind(float px, float py, float pz, float t, int seed, float* out3) {
    v3 w = wind_at(V3(px, py, pz), t, seed);
    out3[0] = w.x; out3[1] = w.y; out3[2] = w.z;
}

int ref_n_presets() { return (int)preset_designs().size(); }
void ref_preset(int i, Genome* out, char* name, int cap) {
    auto d = preset_designs();
    *out = d[i].g;
    snprintf(name, cap, "%s", d[i].name);
}

int ref_sizes(int* genome, int* mass, int* scen, int* state) {
    *genome = sizeof(Genome); *mass = sizeof(MassProps); *scen = sizeof(ScenarioDesc); *state = sizeof(PlaneState);
    return 0;
}

}
