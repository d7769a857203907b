// planes_sim.h -- drop-in replacement for the reference's src/sim.h (sim.h:1-13).
//
// Same seven entry points with the same signatures, semantics and the same Genome /
// float8 byte layouts, implemented over the C ABI of planes_hip.h.  A caller of the
// reference (its main.cu, renderer or any tool) takes this header in place of sim.h and
// links libplanes_hip.so instead of compiling sim.cu.
//
// Include order: the reference's translation units include common.h before sim.h
// (main.cu:1-2, render.cu:1-2), and common.h defines Genome (common.h:60-97).  When it
// has been included (GENOME_N defined, common.h:60), that Genome is the one used here;
// otherwise this header defines the same 140-byte record itself.  float8 belongs to
// sim.h (sim.h:4) and is defined here once.
#pragma once
#include <stdint.h>

#ifndef GENOME_N
#define GENOME_N 35
// Genome: 35 little-endian FP32 genes in this order = the best.bin / pop.bin record
// (common.h:61-97).  Field names follow the reference so existing code compiles.
struct Genome {
    float span, chord, taper, sweep, dihedral_in, dihedral_out, kink, camber1, camber2, washout, elevator,
        nose_mass, keel_h, deck_gap, deck_scale, body_frac, fin_height, fin_chord, fin_sweep, tail_span,
        tail_chord, tail_angle, wing_le_b1, wing_le_b2, chord_b1, chord_b2, fin_layout, fin_cant, fin_taper,
        fin_span_pos, tail_sweep, tail_taper, tail_dihedral, wing_tip_round, center_chord;
};
#endif
static_assert(sizeof(Genome) == 140, "Genome must stay the 140-byte best.bin record");
static_assert(GENOME_N == 35, "Genome must hold the reference's 35 genes");

#ifndef PLANES_SIM_FLOAT8
#define PLANES_SIM_FLOAT8
struct float8 { float a, b, c, d, e, f, g, h; };  // pose frame {pos.xyz, q.wxyz, alive} (sim.cu:121-124)
#endif
static_assert(sizeof(float8) == 32, "float8 must stay the 32-byte pose frame");

void run_ga(int pop, int gens, int nseeds, const char* out_dir, const char* init_path);          // sim.h:6
void evaluate_design(const Genome* g, int scenarios, int base_seed, const char* csv_path);         // sim.h:7
void select_finalist(const Genome* gs, int count, int scenarios, int base_seed, const char* out_dir);  // sim.h:8
void* panelbuf_alloc(int pop);                                                                      // sim.h:9
void panelbuf_free(void* pb);                                                                       // sim.h:10
void gpu_prepare(const Genome* d_gs, void* pb, int pop);                                            // sim.h:11
void gpu_sim(void* pb, int pop, int seed, float* d_fitness, float8* d_rec, int rec_every, int rec_cap,
             float spread, int vary_scenarios);                                                    // sim.h:12-13
