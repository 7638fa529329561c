/*
 * g2switches.h -- read after the user's g2parameters.h.  GABE's parameter
 * file spells its switches as macros (#define calc_gws 1, ...); several of
 * those names are also field names of this library's C structs (gabe_params,
 * gabe_host_init).  Each such macro is read here into a constant of the same
 * name and the macro removed, so the host program uses the names as GABE does
 * and the structs stay usable.  Switches GABE leaves undefined to mean "off"
 * (spec_cut_off, spec_smooth) get their off values.
 */
#ifndef G2SWITCHES_H
#define G2SWITCHES_H

#if fftw_flag == 1
#error "fftw_flag 1 (long double) builds are not supported: the lattice runs in FP64 or FP32 on the GPU"
#endif

#ifndef calc_gws
#error "g2parameters.h must define calc_gws"
#endif
static const int g2sw_calc_gws = calc_gws;
#undef calc_gws
static const int calc_gws = g2sw_calc_gws;

#ifndef expansion_type
#error "g2parameters.h must define expansion_type"
#endif
static const int g2sw_expansion_type = expansion_type;
#undef expansion_type
static const int expansion_type = g2sw_expansion_type;

#ifndef rand_init
#error "g2parameters.h must define rand_init"
#endif
static const int g2sw_rand_init = rand_init;
#undef rand_init
static const int rand_init = g2sw_rand_init;

#ifndef field_full_rand
#error "g2parameters.h must define field_full_rand"
#endif
static const int g2sw_field_full_rand = field_full_rand;
#undef field_full_rand
static const int field_full_rand = g2sw_field_full_rand;

/* undefined spec_cut_off: no window (g2init.cpp:117-130 compile it out) */
#ifdef spec_cut_off
static const double g2sw_spec_cut_off = spec_cut_off;
#undef spec_cut_off
static const double spec_cut_off = g2sw_spec_cut_off;
#else
static const double spec_cut_off = 0.;
#endif
#ifdef spec_smooth
static const double g2sw_spec_smooth = spec_smooth;
#undef spec_smooth
static const double spec_smooth = g2sw_spec_smooth;
#else
static const double spec_smooth = 0.5;
#endif

#ifndef field_outdim
#error "g2parameters.h must define field_outdim"
#endif
static const int g2sw_field_outdim = field_outdim;
#undef field_outdim
static const int field_outdim = g2sw_field_outdim;

#ifndef spec_output
#define spec_output 1
#endif
#ifndef var_output
#define var_output 1
#endif

/* not in GABE's file: the storage form of the lattice (1: FP32 fluctuations
 * about FP64 means, 0: FP64 as GABE) and the checkpoint cadence in steps
 * (0: never); both can be set on the compiler line */
#ifndef GABE_PRECISION
#define GABE_PRECISION 1
#endif
#ifndef GABE_CHECKPOINT_EVERY
#define GABE_CHECKPOINT_EVERY 0
#endif
static const int checkpoint_every = GABE_CHECKPOINT_EVERY;

#endif
