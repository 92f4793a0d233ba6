/* scene.h -- plain-C description of a scene for the ODE step path, and of
 * what a run of it records.  Test infrastructure: shared by the scene driver
 * (an ordinary ODE-API application, compiled against the reference library
 * and against the MI355X library), by the CPU restatement in oracle/ and by
 * the Python test harness (ctypes mirrors these layouts exactly).
 *
 * Creation order is part of the description: bodies, geoms and joints are
 * created in index order, which fixes the reference's list orders
 * (ode.cpp:62-68 prepends bodies/joints, collision_kernel.h:166-171 prepends
 * geoms to the space), and with them pair order, island order and row order.
 */
#ifndef ODEMI_SCENE_H
#define ODEMI_SCENE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { SC_SPHERE = 0, SC_BOX = 1, SC_PLANE = 2, SC_CAPSULE = 3, SC_CYLINDER = 5 };
enum { SC_BALL = 1, SC_HINGE = 2, SC_HINGE2 = 3 };

/* joint_param layout (per joint, SC_NJP doubles) */
enum {
  SCJ_LO = 0, SCJ_HI, SCJ_VEL, SCJ_FMAX, SCJ_FUDGE, SCJ_BOUNCE, SCJ_CFM,
  SCJ_STOP_ERP, SCJ_STOP_CFM, SCJ_SUSP_ERP, SCJ_SUSP_CFM, SCJ_VEL2,
  SCJ_FMAX2, SCJ_ERP, SC_NJP
};

typedef struct SceneDesc {
  /* world */
  double gravity[3];
  double erp, cfm;
  int iterations;
  double sor_w;
  double max_vel, min_depth;     /* contact max correcting vel, surface layer */
  double dt;
  unsigned long seed;            /* dRandSetSeed before the first step */
  int hash_minlevel, hash_maxlevel;
  /* contact policy applied by the near callback */
  int surf_mode;
  double mu, mu2, bounce, bounce_vel, soft_erp, soft_cfm;
  int max_contacts;              /* contacts requested per pair */
  int skip_connected;            /* skip pairs joined by a non-contact joint */
  /* bodies */
  int nb;
  const double *body_pos;        /* nb*3 */
  const double *body_q;          /* nb*4 (w,x,y,z) */
  const double *body_lvel;       /* nb*3 */
  const double *body_avel;       /* nb*3 */
  const double *body_mass;       /* nb   total mass */
  const double *body_I;          /* nb*9 body-frame inertia, row-major */
  /* geoms */
  int ng;
  const int *geom_type;          /* SC_SPHERE / SC_BOX / SC_PLANE */
  const int *geom_body;          /* body index or -1 (static) */
  const double *geom_param;      /* ng*4: radius | lx,ly,lz | a,b,c,d | capsule or cylinder radius, length */
  const double *geom_pos;        /* ng*3 pose of static placeable geoms */
  const double *geom_q;          /* ng*4 */
  /* joints */
  int nj;
  const int *joint_type;         /* SC_BALL / SC_HINGE / SC_HINGE2 */
  const int *joint_b1;           /* body index */
  const int *joint_b2;           /* body index or -1 (static environment) */
  const double *joint_anchor;    /* nj*3 */
  const double *joint_axis1;     /* nj*3 */
  const double *joint_axis2;     /* nj*3 */
  const double *joint_param;     /* nj*SC_NJP, NaN = leave the default */
  /* per-step controls, may be NULL: nsteps*nj*4 = lostop, histop,
     torque1, torque2 (NaN lostop/histop = unchanged) */
  const double *step_ctrl;
  /* may be NULL: per geom 0 or 1.  With a geom in space 1 the run uses two
     top-level hash spaces (space 1 holding static geoms), and every frame
     collides them with dSpaceCollide2(space 0, space 1) before
     dSpaceCollide(space 0), as an application with a static terrain space
     does (demos/raycar/ode.cpp:118-123) */
  const int *geom_space;
  /* 1: step with dWorldStep (the big-matrix LCP solved exactly) instead of
     dWorldQuickStep */
  int world_step;
} SceneDesc;

typedef struct SceneOut {
  int nsteps;
  int record_every;              /* trajectory sample period (steps) */
  double *traj;                  /* (nsteps/record_every+1)*nb*13: pos q lvel avel */
  /* diagnostics for the first diag_steps steps */
  int diag_steps;
  int max_pairs;
  int *pairs;                    /* 2 ints per near-callback call: geom indices */
  int *pair_step_count;          /* diag_steps */
  int max_contacts;
  double *contacts;              /* 9 per contact: pos3 normal3 depth g1 g2 */
  int *contact_step_count;       /* diag_steps */
  double *feedback;              /* 12 per contact joint: f1 t1 f2 t2 */
  /* totals over all steps */
  long long total_pairs, total_contacts;
  double sec_collide, sec_step;  /* wall time in dSpaceCollide / QuickStep+Empty */
  int status;                    /* 0 ok, else failing step + 1 */
  int time_from;                 /* steps before this one are not timed */
  double *residual;              /* may be NULL: nsteps LCP residuals (CPU restatement only) */
  /* may be NULL (MI355X library only): the last step is recorded with
     dWorldSetTape and differentiated with dWorldQuickStepBackward; grad_out
     is dL/d(final state), grad_in receives dL/d(state before the last step),
     nb*13 each (pos 3, q 4, lvel 3, avel 3) */
  const double *grad_out;
  double *grad_in;
} SceneOut;

#ifdef __cplusplus
}
#endif
#endif
