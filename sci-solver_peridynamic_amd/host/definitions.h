// definitions.h -- compile-time constants of the implicit peridynamics path
// (values of the reference's core/definitions.h) and the host vector type.
#ifndef PD_HOST_DEFINITIONS_H
#define PD_HOST_DEFINITIONS_H

#include <stdint.h>

typedef double real_t;

// Host mirror of a particle vector: 4 doubles, w lane unused (core/
// definitions.h:37-39 real4_t = double4).
struct real4_t {
  double x, y, z, w;
};

inline real4_t make_real4(double x, double y, double z, double w) {
  real4_t r = {x, y, z, w};
  return r;
}

#define PD_SMALL (1.0e-10)                          // definitions.h:61
#define SYSTEM_TIMESTEP_BASE (1.0 / 8388608.0)      // definitions.h:65, 2^-23
#define TIMESTEP_PD (128)                           // definitions.h:68
#define TIMESTEP_PD_MIN 100                         // definitions.h:69
#define MAX_SUBSTEP 10                              // definitions.h:71
#define RESTITUTION 0.001                           // definitions.h:74
#define PERIDYNAMICS_MIN_DISTANCE 0.0005            // definitions.h:86
#define PD_V (0.0005 * 0.0005 * 0.0005)             // definitions.h:87
#define PD_K ((5.0) * (2.94315e9))                  // definitions.h:89
#define Kc (1.0e7)                                  // definitions.h:90
#define FCELL (0.75)                                // definitions.h:98
#define MAX_PD_BOND_COUNT 528                       // definitions.h:101
#define PHASEIN (30)                                // definitions.h:103
#define MAX_CG_ITERATION (100000)                   // definitions.h:134
#define CG_RELATIVE_TOLERANCE (1e-20)               // definitions.h:135
#define PD_PI (3.14159265358979323846264)           // definitions.h:62
#define NUM_CELL_X 32                               // definitions.h:54-58
#define NUM_CELL_Y 32
#define NUM_CELL_Z 64
#define NUM_CELLS (NUM_CELL_X * NUM_CELL_Y * NUM_CELL_Z)

enum Integrator { IMPLICIT_EULER = 0, NEWMARK_BETA };

// core/parameters.h:137-144
enum Activity { ACTIVE = 0, SEMI_ACTIVE, INACTIVE, INVISIBLE, NUM_ACTIVITY_MODE };

#endif
