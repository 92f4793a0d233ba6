/* odemi.h -- C ABI of the MI355X-native ODE 0.12 step path.
 *
 * Three groups of entry points:
 *
 *  1. The ODE 0.12 C API subset that the step path sits behind (dReal =
 *     float, i.e. the reference's dSINGLE build).  Signatures, struct layouts
 *     and enum values are those of the reference headers; each declaration
 *     cites the reference declaration it replaces.  The frame loop
 *       dSpaceCollide -> near callback (dCollide, dJointCreateContact,
 *       dJointAttach) -> dWorldQuickStep -> dJointGroupEmpty
 *     runs unchanged against this library: broadphase, narrowphase, row
 *     build, SOR-LCP and integration execute as HIP kernels on the GPU.
 *     dWorldStep (each island's LCP solved exactly), dSpaceCollide2 with
 *     nested spaces, and sphere, box, plane, capsule, cylinder and ray geoms
 *     are part of it.
 *
 *  2. Batch extension (dBatch*): thousands of independent worlds described by
 *     flat arrays, stepped together by one persistent kernel launch per step
 *     (captured in a HIP graph), with a declarative contact policy instead of
 *     a per-pair host callback.
 *
 *  3. Adjoint: dBatchStepBackward / dWorldQuickStepBackward, the hand-written
 *     reverse-mode derivative of one step.
 *
 * Errors follow the reference's convention (no exceptions, no errno):
 * counts are returned as int, dWorldQuickStep returns 1 on success and 0 on
 * failure with the world unchanged (include/ode/objects.h:342-348), user
 * errors go to the handler installed with dSetErrorHandler / dSetDebugHandler.
 */
#ifndef ODEMI_H
#define ODEMI_H

#include <math.h>
#include <stdarg.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define dSINGLE 1
typedef float dReal;
#ifndef ODE_API
#define ODE_API
#endif

/* include/ode/common.h:139-144 */
typedef dReal dVector3[4];
typedef dReal dVector4[4];
typedef dReal dMatrix3[4 * 3];
typedef dReal dMatrix4[4 * 4];
typedef dReal dMatrix6[8 * 6];
typedef dReal dQuaternion[4];

#define dInfinity INFINITY

struct dxWorld;
struct dxSpace;
struct dxBody;
struct dxGeom;
struct dxJoint;
struct dxJointGroup;
typedef struct dxWorld *dWorldID;
typedef struct dxSpace *dSpaceID;
typedef struct dxBody *dBodyID;
typedef struct dxGeom *dGeomID;
typedef struct dxJoint *dJointID;
typedef struct dxJointGroup *dJointGroupID;

/* joint types: include/ode/common.h:250-267 */
typedef enum {
  dJointTypeNone = 0, dJointTypeBall, dJointTypeHinge, dJointTypeSlider, dJointTypeContact,
  dJointTypeUniversal, dJointTypeHinge2, dJointTypeFixed, dJointTypeNull, dJointTypeAMotor,
  dJointTypeLMotor, dJointTypePlane2D, dJointTypePR, dJointTypePU, dJointTypePiston
} dJointType;

/* joint parameters: include/ode/common.h:303-355 */
enum {
  dParamLoStop = 0, dParamHiStop, dParamVel, dParamFMax, dParamFudgeFactor, dParamBounce, dParamCFM,
  dParamStopERP, dParamStopCFM, dParamSuspensionERP, dParamSuspensionCFM, dParamERP,
  dParamsInGroup,
  dParamGroup1 = 0x000, dParamLoStop1 = 0x000, dParamHiStop1, dParamVel1, dParamFMax1, dParamFudgeFactor1,
  dParamBounce1, dParamCFM1, dParamStopERP1, dParamStopCFM1, dParamSuspensionERP1, dParamSuspensionCFM1, dParamERP1,
  dParamGroup2 = 0x100, dParamLoStop2 = 0x100, dParamHiStop2, dParamVel2, dParamFMax2, dParamFudgeFactor2,
  dParamBounce2, dParamCFM2, dParamStopERP2, dParamStopCFM2, dParamSuspensionERP2, dParamSuspensionCFM2, dParamERP2,
  dParamGroup = 0x100
};

/* geom classes: include/ode/collision.h:880-902 */
enum {
  dSphereClass = 0, dBoxClass, dCapsuleClass, dCylinderClass, dPlaneClass, dRayClass, dConvexClass,
  dGeomTransformClass, dTriMeshClass, dHeightfieldClass,
  dFirstSpaceClass, dSimpleSpaceClass = dFirstSpaceClass, dHashSpaceClass, dSweepAndPruneSpaceClass,
  dQuadTreeSpaceClass, dLastSpaceClass = dQuadTreeSpaceClass,
  dFirstUserClass, dLastUserClass = dFirstUserClass + 4 - 1, dGeomNumClasses
};

/* contact surface modes: include/ode/contact.h:33-49 */
enum {
  dContactMu2 = 0x001, dContactFDir1 = 0x002, dContactBounce = 0x004, dContactSoftERP = 0x008,
  dContactSoftCFM = 0x010, dContactMotion1 = 0x020, dContactMotion2 = 0x040, dContactMotionN = 0x080,
  dContactSlip1 = 0x100, dContactSlip2 = 0x200,
  dContactApprox0 = 0x0000, dContactApprox1_1 = 0x1000, dContactApprox1_2 = 0x2000, dContactApprox1 = 0x3000
};

/* include/ode/contact.h:52-65 */
typedef struct dSurfaceParameters {
  int mode;
  dReal mu;
  dReal mu2;
  dReal bounce;
  dReal bounce_vel;
  dReal soft_erp;
  dReal soft_cfm;
  dReal motion1, motion2, motionN;
  dReal slip1, slip2;
} dSurfaceParameters;

/* include/ode/contact.h:82-88: the normal points into body 1 */
typedef struct dContactGeom {
  dVector3 pos;
  dVector3 normal;
  dReal depth;
  dGeomID g1, g2;
  int side1, side2;
} dContactGeom;

/* include/ode/contact.h:93-97 */
typedef struct dContact {
  dSurfaceParameters surface;
  dContactGeom geom;
  dVector3 fdir1;
} dContact;

/* include/ode/mass.h:88-92 */
typedef struct dMass {
  dReal mass;
  dVector3 c;
  dMatrix3 I;
} dMass;

/* include/ode/common.h:368-373 */
typedef struct dJointFeedback {
  dVector3 f1, t1, f2, t2;
} dJointFeedback;

/* include/ode/collision_space.h:49 */
typedef void dNearCallback(void *data, dGeomID o1, dGeomID o2);

/* include/ode/error.h */
typedef void dMessageFunction(int errnum, const char *msg, va_list ap);

/* ------------------------------------------------------------------ 1. ODE API */

/* init: include/ode/odeinit.h:119, 227 */
ODE_API int dInitODE2(unsigned int uiInitFlags);
ODE_API void dInitODE(void);
ODE_API void dCloseODE(void);

/* errors: include/ode/error.h (ode/src/error.cpp:33-45) */
ODE_API void dSetErrorHandler(dMessageFunction *fn);
ODE_API void dSetDebugHandler(dMessageFunction *fn);
ODE_API void dSetMessageHandler(dMessageFunction *fn);

/* world: include/ode/objects.h:54-440 */
ODE_API dWorldID dWorldCreate(void);                                          /* objects.h:54 */
ODE_API void dWorldDestroy(dWorldID w);                                       /* objects.h:66 */
ODE_API void dWorldSetGravity(dWorldID w, dReal x, dReal y, dReal z);         /* objects.h:77 */
ODE_API void dWorldGetGravity(dWorldID w, dVector3 gravity);
ODE_API void dWorldSetERP(dWorldID w, dReal erp);                             /* objects.h:94 */
ODE_API dReal dWorldGetERP(dWorldID w);
ODE_API void dWorldSetCFM(dWorldID w, dReal cfm);                             /* objects.h:111 */
ODE_API dReal dWorldGetCFM(dWorldID w);
ODE_API void dWorldSetQuickStepNumIterations(dWorldID w, int num);            /* objects.h:385 */
ODE_API int dWorldGetQuickStepNumIterations(dWorldID w);
ODE_API void dWorldSetQuickStepW(dWorldID w, dReal over_relaxation);          /* objects.h:401 */
ODE_API dReal dWorldGetQuickStepW(dWorldID w);
ODE_API void dWorldSetContactMaxCorrectingVel(dWorldID w, dReal vel);         /* objects.h:420 */
ODE_API void dWorldSetContactSurfaceLayer(dWorldID w, dReal depth);           /* objects.h:440 */
ODE_API int dWorldQuickStep(dWorldID w, dReal stepsize);                      /* objects.h:352 */
/* dWorldStep: the same step with each island's LCP solved exactly instead of
 * by the 20 SOR iterations -- the answer dSolveLCP (lcp.cpp) defines: the
 * rows without a friction index solved first, then the friction bounds fixed
 * from those normal impulses and all rows solved (block principal pivoting
 * on a Cholesky factorisation of the free rows, one workgroup per world).
 * The workspace holds an island of every row of the world (2 m^2 floats for
 * m rows).  Returns 0 with an error when the solve does not finish. */
ODE_API int dWorldStep(dWorldID w, dReal stepsize);                           /* objects.h:309 */

/* bodies: include/ode/objects.h:856-1075 */
ODE_API dBodyID dBodyCreate(dWorldID w);                                      /* objects.h:856 */
ODE_API void dBodyDestroy(dBodyID b);                                         /* objects.h:866 */
ODE_API void dBodySetData(dBodyID b, void *data);                             /* objects.h:873 */
ODE_API void *dBodyGetData(dBodyID b);                                        /* objects.h:880 */
ODE_API void dBodySetPosition(dBodyID b, dReal x, dReal y, dReal z);          /* objects.h:890 */
ODE_API void dBodySetRotation(dBodyID b, const dMatrix3 R);                   /* objects.h:900 */
ODE_API void dBodySetQuaternion(dBodyID b, const dQuaternion q);              /* objects.h:910 */
ODE_API void dBodySetLinearVel(dBodyID b, dReal x, dReal y, dReal z);         /* objects.h:916 */
ODE_API void dBodySetAngularVel(dBodyID b, dReal x, dReal y, dReal z);        /* objects.h:922 */
ODE_API const dReal *dBodyGetPosition(dBodyID b);                             /* objects.h:933 */
ODE_API const dReal *dBodyGetRotation(dBodyID b);                             /* objects.h:951 */
ODE_API const dReal *dBodyGetQuaternion(dBodyID b);                           /* objects.h:969 */
ODE_API const dReal *dBodyGetLinearVel(dBodyID b);                            /* objects.h:986 */
ODE_API const dReal *dBodyGetAngularVel(dBodyID b);                           /* objects.h:992 */
ODE_API void dBodySetMass(dBodyID b, const dMass *mass);                      /* objects.h:998 */
ODE_API void dBodyGetMass(dBodyID b, dMass *mass);                            /* objects.h:1004 */
ODE_API void dBodyAddForce(dBodyID b, dReal fx, dReal fy, dReal fz);          /* objects.h:1010 */
ODE_API void dBodyAddTorque(dBodyID b, dReal fx, dReal fy, dReal fz);         /* objects.h:1016 */
ODE_API const dReal *dBodyGetForce(dBodyID b);                                /* objects.h:1064 */
ODE_API const dReal *dBodyGetTorque(dBodyID b);                               /* objects.h:1075 */
ODE_API dWorldID dBodyGetWorld(dBodyID b);

/* mass: include/ode/mass.h:45-74 */
ODE_API void dMassSetZero(dMass *m);                                          /* mass.h:45 */
ODE_API void dMassSetParameters(dMass *m, dReal themass, dReal cgx, dReal cgy, dReal cgz, dReal I11,
                                dReal I22, dReal I33, dReal I12, dReal I13, dReal I23); /* mass.h:47 */
ODE_API void dMassSetSphere(dMass *m, dReal density, dReal radius);           /* mass.h:52 */
ODE_API void dMassSetSphereTotal(dMass *m, dReal total_mass, dReal radius);   /* mass.h:53 */
ODE_API void dMassSetBox(dMass *m, dReal density, dReal lx, dReal ly, dReal lz);          /* mass.h:65 */
ODE_API void dMassSetBoxTotal(dMass *m, dReal total_mass, dReal lx, dReal ly, dReal lz);  /* mass.h:67 */
ODE_API void dMassAdjust(dMass *m, dReal newmass);                            /* mass.h:74 */
ODE_API void dMassSetCapsule(dMass *m, dReal density, int direction, dReal radius, dReal length);      /* mass.h:55 */
ODE_API void dMassSetCapsuleTotal(dMass *m, dReal total_mass, int direction, dReal radius, dReal length); /* mass.h:57 */
ODE_API void dMassSetCylinder(dMass *m, dReal density, int direction, dReal radius, dReal length);     /* mass.h:60 */
ODE_API void dMassSetCylinderTotal(dMass *m, dReal total_mass, int direction, dReal radius, dReal length); /* mass.h:62 */

/* random numbers: include/ode/misc.h:42-51 */
ODE_API unsigned long dRand(void);                                            /* misc.h:42 */
ODE_API unsigned long dRandGetSeed(void);                                     /* misc.h:45 */
ODE_API void dRandSetSeed(unsigned long s);                                   /* misc.h:46 */
ODE_API int dRandInt(int n);                                                  /* misc.h:51 */
ODE_API int dTestRand(void);

/* joints: include/ode/objects.h:1538-1916 */
ODE_API dJointGroupID dJointGroupCreate(int max_size);                        /* objects.h:1655 */
ODE_API void dJointGroupDestroy(dJointGroupID g);                             /* objects.h:1663 */
ODE_API void dJointGroupEmpty(dJointGroupID g);                               /* objects.h:1672 */
ODE_API dJointID dJointCreateBall(dWorldID w, dJointGroupID g);               /* objects.h:1538 */
ODE_API dJointID dJointCreateHinge(dWorldID w, dJointGroupID g);              /* objects.h:1546 */
ODE_API dJointID dJointCreateContact(dWorldID w, dJointGroupID g, const dContact *c); /* objects.h:1562 */
ODE_API dJointID dJointCreateHinge2(dWorldID w, dJointGroupID g);             /* objects.h:1570 */
ODE_API void dJointDestroy(dJointID j);                                       /* objects.h:1647 */
ODE_API void dJointAttach(dJointID j, dBodyID b1, dBodyID b2);                /* objects.h:1693 */
ODE_API int dJointGetType(dJointID j);                                        /* objects.h:1749 */
ODE_API dBodyID dJointGetBody(dJointID j, int index);                         /* objects.h:1761 */
ODE_API void dJointSetFeedback(dJointID j, dJointFeedback *fb);               /* objects.h:1770 */
ODE_API dJointFeedback *dJointGetFeedback(dJointID j);                        /* objects.h:1776 */
ODE_API void dJointSetBallAnchor(dJointID j, dReal x, dReal y, dReal z);      /* objects.h:1785 */
ODE_API void dJointSetBallParam(dJointID j, int parameter, dReal value);      /* objects.h:1797 */
ODE_API void dJointSetHingeAnchor(dJointID j, dReal x, dReal y, dReal z);     /* objects.h:1803 */
ODE_API void dJointSetHingeAxis(dJointID j, dReal x, dReal y, dReal z);       /* objects.h:1811 */
ODE_API void dJointSetHingeParam(dJointID j, int parameter, dReal value);     /* objects.h:1847 */
ODE_API void dJointSetHinge2Anchor(dJointID j, dReal x, dReal y, dReal z);    /* objects.h:1890 */
ODE_API void dJointSetHinge2Axis1(dJointID j, dReal x, dReal y, dReal z);     /* objects.h:1896 */
ODE_API void dJointSetHinge2Axis2(dJointID j, dReal x, dReal y, dReal z);     /* objects.h:1902 */
ODE_API void dJointSetHinge2Param(dJointID j, int parameter, dReal value);    /* objects.h:1908 */
ODE_API void dJointAddHinge2Torques(dJointID j, dReal torque1, dReal torque2); /* objects.h:1916 */
ODE_API int dAreConnected(dBodyID a, dBodyID b);                              /* objects.h:2953 */
ODE_API int dAreConnectedExcluding(dBodyID a, dBodyID b, int joint_type);     /* objects.h:2968 */

/* spaces: include/ode/collision_space.h:52-153 */
ODE_API dSpaceID dSimpleSpaceCreate(dSpaceID space);                          /* collision_space.h:52 */
ODE_API dSpaceID dHashSpaceCreate(dSpaceID space);                            /* collision_space.h:53 */
ODE_API void dSpaceDestroy(dSpaceID space);                                   /* collision_space.h:70 */
ODE_API void dHashSpaceSetLevels(dSpaceID space, int minlevel, int maxlevel); /* collision_space.h:72 */
ODE_API void dSpaceAdd(dSpaceID space, dGeomID g);                            /* collision_space.h:149 */
ODE_API void dSpaceRemove(dSpaceID space, dGeomID g);                         /* collision_space.h:150 */
ODE_API int dSpaceGetNumGeoms(dSpaceID space);                                /* collision_space.h:153 */
ODE_API dGeomID dSpaceGetGeom(dSpaceID space, int i);                         /* collision_space.h:154 */
ODE_API int dSpaceQuery(dSpaceID space, dGeomID g);                           /* collision_space.h:151 */
ODE_API void dSpaceSetCleanup(dSpaceID space, int mode);                      /* collision_space.h:75 */
ODE_API int dSpaceGetCleanup(dSpaceID space);                                 /* collision_space.h:76 */
ODE_API void dSpaceSetSublevel(dSpaceID space, int sublevel);                 /* collision_space.h:100 */
ODE_API int dSpaceGetSublevel(dSpaceID space);                                /* collision_space.h:114 */
ODE_API int dGeomIsSpace(dGeomID geom);                                       /* collision.h:273 */
ODE_API dSpaceID dGeomGetSpace(dGeomID geom);                                 /* collision.h:283 */

/* collision: include/ode/collision.h:68-1048 */
ODE_API void dSpaceCollide(dSpaceID space, void *data, dNearCallback *callback); /* collision.h:825 */
/* Spaces are geoms: a space created with a parent sits in that space, and a
 * dSpaceID is passed where a dGeomID is taken by a cast, as in the reference.
 * dSpaceCollide on a space that holds spaces is refused in this build (the
 * error handler says so); dSpaceCollide2 takes spaces at any level. */
ODE_API void dSpaceCollide2(dGeomID space1, dGeomID space2, void *data, dNearCallback *callback); /* collision.h:868 */
/* flags bit: any contacts will do (collision.h:746); the colliders here
 * refine nothing, so it changes no result */
#define CONTACTS_UNIMPORTANT 0x80000000
ODE_API int dCollide(dGeomID o1, dGeomID o2, int flags, dContactGeom *contact, int skip); /* collision.h:795 */
/* No counterpart in the reference.  dCollide for callers that cannot
 * allocate device memory (code inside a serving loop or a harness that
 * forbids allocation): the same contacts as dCollide, computed in a
 * library-owned scratch array of one pair's contacts per GPU (set up by
 * dInitODE2 or on the GPU's first call), on the calling thread's per-thread
 * stream.  At most 8 contacts may be requested (flags & 0xffff <= 8, the most
 * any collider here returns); a larger request is refused with an error and
 * returns 0.  Calls that overlap on one GPU share the array and race. */
ODE_API int dCollideNoAlloc(dGeomID o1, dGeomID o2, int flags, dContactGeom *contact, int skip);
ODE_API dGeomID dCreateSphere(dSpaceID space, dReal radius);                  /* collision.h:924 */
ODE_API dGeomID dCreateBox(dSpaceID space, dReal lx, dReal ly, dReal lz);     /* collision.h:1002 */
ODE_API dGeomID dCreatePlane(dSpaceID space, dReal a, dReal b, dReal c, dReal d); /* collision.h:1046 */
ODE_API dGeomID dCreateCapsule(dSpaceID space, dReal radius, dReal length);  /* collision.h:1051 */
ODE_API void dGeomCapsuleSetParams(dGeomID ccylinder, dReal radius, dReal length); /* collision.h:1052 */
ODE_API void dGeomCapsuleGetParams(dGeomID ccylinder, dReal *radius, dReal *length); /* collision.h:1053 */
/* Cylinders (cylinder.cpp:47-103): the axis is column 2 of the rotation.
 * Their colliders are those of the reference without libccd: cylinder-sphere,
 * cylinder-plane, cylinder-box and ray-cylinder; a cylinder against a
 * cylinder or a capsule gives no contacts, as there. */
ODE_API dGeomID dCreateCylinder(dSpaceID space, dReal radius, dReal length);  /* collision.h:1063 */
ODE_API void dGeomCylinderSetParams(dGeomID cylinder, dReal radius, dReal length); /* collision.h:1064 */
ODE_API void dGeomCylinderGetParams(dGeomID cylinder, dReal *radius, dReal *length); /* collision.h:1065 */
/* Rays (ray.cpp:47-176): a ray starts at its geom position and points along
 * column 2 of its rotation; its contacts have depth = distance from the start
 * and the normal of the surface hit (reversed when the ray starts inside).
 * A ray attached to a body follows the body; dGeomRaySet on such a ray moves
 * the body (its position) and is refused for a direction other than the
 * body's own z axis, since the body's rotation here comes from its
 * quaternion.  The FirstContact / BackfaceCull / ClosestHit flags are stored
 * and returned; no collider here reads them (in the reference only the
 * trimesh and heightfield colliders do). */
ODE_API dGeomID dCreateRay(dSpaceID space, dReal length);                     /* collision.h:1067 */
ODE_API void dGeomRaySetLength(dGeomID ray, dReal length);                    /* collision.h:1068 */
ODE_API dReal dGeomRayGetLength(dGeomID ray);                                 /* collision.h:1069 */
ODE_API void dGeomRaySet(dGeomID ray, dReal px, dReal py, dReal pz, dReal dx, dReal dy, dReal dz); /* collision.h:1070 */
ODE_API void dGeomRayGet(dGeomID ray, dVector3 start, dVector3 dir);          /* collision.h:1072 */
ODE_API void dGeomRaySetParams(dGeomID g, int FirstContact, int BackfaceCull); /* collision.h:1079 */
ODE_API void dGeomRayGetParams(dGeomID g, int *FirstContact, int *BackfaceCull); /* collision.h:1080 */
ODE_API void dGeomRaySetClosestHit(dGeomID g, int closestHit);                /* collision.h:1081 */
ODE_API int dGeomRayGetClosestHit(dGeomID g);                                 /* collision.h:1082 */
ODE_API void dGeomDestroy(dGeomID g);                                         /* collision.h:68 */
ODE_API void dGeomSetData(dGeomID g, void *data);                             /* collision.h:78 */
ODE_API void *dGeomGetData(dGeomID g);                                        /* collision.h:87 */
ODE_API void dGeomSetBody(dGeomID g, dBodyID b);                              /* collision.h:108 */
ODE_API dBodyID dGeomGetBody(dGeomID g);                                      /* collision.h:117 */
ODE_API void dGeomSetPosition(dGeomID g, dReal x, dReal y, dReal z);          /* collision.h:134 */
ODE_API void dGeomSetRotation(dGeomID g, const dMatrix3 R);                   /* collision.h:149 */
ODE_API void dGeomSetQuaternion(dGeomID g, const dQuaternion q);              /* collision.h:165 */
ODE_API const dReal *dGeomGetPosition(dGeomID g);                             /* collision.h:184 */
ODE_API const dReal *dGeomGetRotation(dGeomID g);                             /* collision.h:213 */
ODE_API void dGeomGetAABB(dGeomID g, dReal aabb[6]);                          /* collision.h:264 */
ODE_API int dGeomGetClass(dGeomID g);                                         /* collision.h:310 */
ODE_API void dGeomSetCategoryBits(dGeomID g, unsigned long bits);             /* collision.h:325 */
ODE_API void dGeomSetCollideBits(dGeomID g, unsigned long bits);              /* collision.h:340 */
ODE_API dReal dGeomSphereGetRadius(dGeomID g);                                /* collision.h:947 */
ODE_API void dGeomBoxGetLengths(dGeomID g, dVector3 result);                  /* collision.h:1028 */
ODE_API void dGeomPlaneGetParams(dGeomID g, dVector4 result);                 /* collision.h:1048 */

/* ------------------------------------------------------------------ 2. batch */

/* Step order of the SOR-LCP solver.
 *  ODEMI_ORDER_REFERENCE: the reference's islands, row order and seeded row
 *    shuffle (quickstep.cpp:406-482, misc.cpp:33-117), pairs in the hash
 *    space's callback order; results track the reference build.
 *  ODEMI_ORDER_COLORED: graph-coloured projected Gauss-Seidel (constraints of
 *    one colour share no body and are relaxed in parallel; a colour holds at
 *    most one workgroup width of constraints, so results are reproducible at
 *    one launch shape and differ in rounding between shapes). */
enum { ODEMI_ORDER_REFERENCE = 0, ODEMI_ORDER_COLORED = 1 };

/* one world of a batch; meaning as the corresponding ODE setters */
typedef struct dBatchWorldDesc {
  double gravity[3];
  double erp, cfm;
  int iterations;
  double sor_w;
  double max_vel, min_depth;
  double dt;
  unsigned long seed;            /* SOR shuffle seed (dRandSetSeed) */
  int hash_minlevel, hash_maxlevel;
  /* contact policy applied to every pair (replaces the near callback) */
  int surf_mode;
  double mu, mu2, bounce, bounce_vel, soft_erp, soft_cfm;
  int max_contacts;              /* contacts per pair, 1..8 */
  int skip_connected;            /* skip pairs joined by a non-contact joint */
  int nb, ng, nj;
} dBatchWorldDesc;

/* Flat arrays of all worlds, world after world; body/geom/joint indices are
 * local to their world.  Layout per element as in dBatchWorldDesc's
 * comments: body_pos 3, body_q 4 (w,x,y,z), body_lvel 3, body_avel 3,
 * body_mass 1, body_I 9 (body frame), geom_param 4 (radius | lx,ly,lz |
 * a,b,c,d), geom_pos 3 and geom_q 4 (static geoms), joint_anchor 3,
 * joint_axis1 3, joint_axis2 3, joint_param ODEMI_NJP (NaN = default). */
#define ODEMI_NJP 14
enum { ODEMI_GEOM_SPHERE = 0, ODEMI_GEOM_BOX = 1, ODEMI_GEOM_PLANE = 2 };
enum { ODEMI_JOINT_BALL = 1, ODEMI_JOINT_HINGE = 2, ODEMI_JOINT_HINGE2 = 3 };
typedef struct dBatchDesc {
  int nworlds;
  const dBatchWorldDesc *worlds;
  const double *body_pos, *body_q, *body_lvel, *body_avel, *body_mass, *body_I;
  const int *geom_type, *geom_body;
  const double *geom_param, *geom_pos, *geom_q;
  const int *joint_type, *joint_b1, *joint_b2;
  const double *joint_anchor, *joint_axis1, *joint_axis2, *joint_param;
} dBatchDesc;

struct dxBatch;
typedef struct dxBatch *dBatchID;

/* Creates device state for all worlds (0 on failure). */
ODE_API dBatchID dBatchCreate(const dBatchDesc *desc, int order_mode);
ODE_API void dBatchDestroy(dBatchID b);
/* Batched path: one persistent launch per step for all worlds, replayed from
 * a HIP graph.  Returns 1 on success, 0 on failure (state unchanged). */
ODE_API int dBatchStep(dBatchID b, int nsteps);
/* dBatchStep bracketed by HIP events on the batch's stream: *ms receives
 * the device time of the nsteps launches. */
ODE_API int dBatchStepTimed(dBatchID b, int nsteps, float *ms);
/* Plain path: the same step, one world per launch (collide + step kernels
 * per world), no graph. */
ODE_API int dBatchStepPlain(dBatchID b, int nsteps);
/* Per-joint controls applied at the start of every following step:
 * nj_total*4 doubles (lostop, histop, torque1, torque2; NaN = unchanged). */
ODE_API int dBatchSetControls(dBatchID b, const double *ctrl);
/* Body state, 13 floats per body: pos 3, q 4, lvel 3, avel 3. */
ODE_API int dBatchGetState(dBatchID b, float *out);
ODE_API int dBatchSetState(dBatchID b, const float *in);
/* Totals of the last step: [pairs, contacts, rows, islands_or_colours, overflow]. */
ODE_API int dBatchGetCounters(dBatchID b, long long *out5);
/* Counters for every world, 8 ints per world.  Of the last step: pairs,
 * contacts, rows, islands or colours, overflow bits (1 grid cells, 2 pairs,
 * 4 contacts, 8 rows, 16 non-finite state: the world is frozen), SOR
 * iterations run.  Since dBatchResetSolveStats (or dBatchCreate): the number
 * of iterate-to-tolerance solves that ended at the iteration cap with the
 * residual above the tolerance, and the largest final residual of such a
 * solve (the bits of a float). */
ODE_API int dBatchGetWorldCounters(dBatchID b, int *out);
/* Zeroes the two accumulating counters of dBatchGetWorldCounters. */
ODE_API int dBatchResetSolveStats(dBatchID b);
/* Coloured order only: iterate each world's SOR until its LCP residual is
 * at most tol (checked every check_every iterations, at most max_iterations);
 * tol = 0 restores the fixed iteration count of the world.  The reference
 * order always runs the fixed count, as dWorldQuickStep does. */
ODE_API int dBatchSetSolverTolerance(dBatchID b, float tol, int max_iterations, int check_every);
/* As dBatchSetSolverTolerance; accel = 1 adds a nonsmooth nonlinear
 * conjugate-gradient step after every sweep (same fixed point, fewer sweeps
 * on stacked contacts).  A world whose state becomes non-finite is frozen
 * and flagged (overflow bit 16) until dBatchSetState. */
ODE_API int dBatchSetSolverToleranceAccel(dBatchID b, float tol, int max_iterations, int check_every, int accel);
/* Coloured order only: warm start.  Each solve starts from warm (0..1) times
 * the impulses of the previous step -- contacts matched by geom pair and
 * index within the pair, permanent joints by joint when their row count is
 * unchanged -- instead of from zero; the solution at a tolerance is the same
 * LCP's.  The reference keeps this behind WARM_STARTING for joints only
 * (ode/src/quickstep.cpp:43-47, 348-372, 864-895), as contacts are
 * recreated every step.  dBatchSetState and this call clear the stored
 * impulses.  warm = 0 (the default) restores the reference's zero start. */
ODE_API int dBatchSetWarmStart(dBatchID b, float warm);
/* dWorldStep for every world of a batch (no counterpart in the reference's
 * batch-less API; the single-world form is dWorldStep): each island's LCP
 * solved exactly instead of by SOR iterations, in the reference order
 * (ODEMI_ORDER_REFERENCE; refused for the coloured order).  Each world gets
 * a device workspace for islands of up to max_island_rows rows
 * (4 (2 m^2 + 7 m) bytes, m = max_island_rows); a larger island flags
 * OVF_LCP (32) in the world's overflow counter and keeps the clamped last
 * iterate.  0 switches back to the SOR iterations and frees the workspaces. */
ODE_API int dBatchSetExactLCP(dBatchID b, int max_island_rows);
/* Coloured order at a tolerance: on = 1 runs each step as three kernels --
 * collide and the solver set-up, the sweeps, the integration -- so that the
 * sweeps (the bulk of the work at a tolerance) get a lean kernel with its own
 * workgroup size wg, LDS pool lds (bytes) and persistent grid (0: one
 * workgroup per world).  Results are those of the one-kernel step. */
ODE_API int dBatchSetSorLaunch(dBatchID b, int on, int wg, int lds, int grid);
ODE_API int dBatchGetSorLaunch(dBatchID b, int *on, int *wg, int *lds, int *grid);
/* Enables the LCP complementarity residual of the last step per world
 * (max_i |lambda_i - clamp(lambda_i - (A lambda - b)_i / A_ii)|). */
ODE_API int dBatchSetResidualTracking(dBatchID b, int on);
ODE_API int dBatchGetResidual(dBatchID b, float *out_per_world);
/* Chooses launch shape and grid cell scale by timing candidates on a copy
 * of the state; returns the chosen workgroup size. */
ODE_API int dBatchAutotune(dBatchID b, int trial_steps);
/* dBatchAutotune with a reduced search: ODEMI_TUNE_KEEP_CELL keeps the grid
 * cell scale, ODEMI_TUNE_FUSED_ONLY tries only the fused kernel (while the
 * two-kernel step is switched off, every search is fused only).  The trials
 * run in the batch's current solver mode. */
#define ODEMI_TUNE_KEEP_CELL 1
#define ODEMI_TUNE_FUSED_ONLY 2
#define ODEMI_TUNE_NO_SOR_SPLIT 4   /* do not try the sweeps in a kernel of their own */
ODE_API int dBatchAutotuneEx(dBatchID b, int trial_steps, int flags);
ODE_API int dBatchGetLaunchConfig(dBatchID b, int *wg_size, float *cell_scale);
ODE_API int dBatchSetLaunchConfig(dBatchID b, int wg_size, float cell_scale);
/* LDS pool per workgroup in bytes (sort keys, colouring state and the
 * coloured solver's working set live there when they fit; larger worlds
 * use their device scratch).  Set returns 0 for a size above the device's
 * per-workgroup limit. */
ODE_API int dBatchGetLds(dBatchID b);
/* 1: a step runs as two kernels (collision, then quickstep), each with its
 * own register budget; 0: one fused kernel.  The two-kernel step is switched
 * off for now (a 4096-world timing run with it did not finish, cause not
 * found): dBatchSetSplit(b, 1) returns 0 and the autotuner keeps the fused
 * kernel. */
ODE_API int dBatchGetSplit(dBatchID b);
/* Persistent grid size (workgroups pulling worlds from a queue); 0 launches
 * one workgroup per world.  dBatchGetLaunchConfig returns the current one. */
ODE_API int dBatchSetPersistentGrid(dBatchID b, int grid);
/* Order in which the persistent grid pulls worlds: 0 (default) by size
 * (bodies + geoms), largest first; 1 by estimated solver work (the larger of
 * the last step's rows x sweeps and 3/4 of the previous estimate), most
 * expensive first, up to 8192 worlds (larger batches keep the size order).
 * On the uneven-worlds workload at a tolerance the size order is the faster
 * one (198 vs 204 ms per step).  No counterpart in the reference. */
ODE_API int dBatchSetQueueOrder(dBatchID b, int by_cost);
ODE_API int dBatchGetQueueOrder(dBatchID b);
ODE_API int dBatchSetSplit(dBatchID b, int on);
ODE_API int dBatchSetLds(dBatchID b, int bytes);
/* In-kernel phase clocks: per phase (AABB, broadphase, narrowphase, body
 * setup + getInfo1, getInfo2 + rhs, ordering/colouring, SOR, integration)
 * the workgroup time summed over worlds since the last read, in ms (8
 * doubles).  Reading resets the clocks. */
ODE_API int dBatchSetProfiling(dBatchID b, int on);
ODE_API int dBatchGetPhaseTimes(dBatchID b, double *ms8);
/* The same per world (nworlds x 8 values, in dBatchGetPhaseTimes' phase
 * order), also read-and-clear. */
ODE_API int dBatchGetWorldPhaseTimes(dBatchID b, double *ms);
/* Device pointer of the body positions (one float4 per body, w unused) and
 * the HIP stream the batch runs on, for zero-copy use from other GPU code. */
/* Diagnostics of the last step of one world.  dBatchGetPairs writes the
 * candidate pairs (2 world-local geom ids each) in report order;
 * dBatchGetContacts writes 12 floats per contact: pos 3, normal 3 (into
 * body 1), depth, g1, g2, and the impulses of its up to 3 rows (normal,
 * friction 1, friction 2).  Both return the full count of the step (at
 * most maxn entries are written), -1 for a bad world index. */
ODE_API int dBatchGetPairs(dBatchID b, int world, int maxn, int *out);
ODE_API int dBatchGetContacts(dBatchID b, int world, int maxn, float *out);
ODE_API void *dBatchDeviceState(dBatchID b);
ODE_API void *dBatchStream(dBatchID b);
ODE_API int dBatchNumBodies(dBatchID b);

/* ------------------------------------------------------------------ 3. adjoint */

/* Starts recording the steps run by dBatchStep (at most `max_steps`): the
 * state is checkpointed every checkpoint_interval steps (default 10); the
 * backward pass replays each segment from its checkpoint (with the joint
 * controls each step ran with).  Worlds may hold spheres, boxes and planes,
 * ball joints, hinges without limits, hinge2 joints (suspension, steering
 * limits, motors, dBatchSetControls torques), with dContactApprox1 friction
 * (or mu = 0) and optional dContactBounce; a recorded step with a row the
 * adjoint does not cover (a bouncy or motor-loaded one-sided limit) makes
 * dBatchStepBackward fail with an error.  The
 * derivative is that of the converged LCP: run the coloured order with a
 * solver tolerance (dBatchSetSolverTolerance) for accurate gradients. */
ODE_API int dBatchTapeBegin(dBatchID b, int max_steps);
ODE_API int dBatchTapeEnd(dBatchID b);
/* Checkpoint interval, active-set threshold on impulses, relative residual
 * and iteration cap of the adjoint linear solve (not while recording). */
ODE_API int dBatchTapeSetOptions(dBatchID b, int checkpoint_interval, float act_tol, float lin_tol, int lin_maxit);
/* on = 1 (before dBatchTapeBegin): the forward pass also keeps every
 * recorded step's pre-state and contacts on the device (64 bytes per body
 * plus the contact capacity, per step), and dBatchStepBackward runs the
 * reverse sweep without replaying the checkpoint segments; the gradients
 * are the same bits.  No counterpart in the reference (which has no
 * adjoint). */
ODE_API int dBatchTapeKeepSteps(dBatchID b, int on);
/* Totals of the last backward pass: adjoint-solve iterations, and contact
 * pairs whose dual-number collider disagreed with the tape (0 expected). */
ODE_API int dBatchTapeStats(dBatchID b, long long *out2);
/* Reverse-mode derivative of the last `nsteps` recorded steps.
 * grad_out: dL/d(state after), 13 floats per body (pos, q, lvel, avel).
 * grad_in:  receives dL/d(state before the first of those steps).
 * grad_mu:  receives dL/d(mu) per world (may be NULL).
 * grad_bounce: receives dL/d(bounce) per world (may be NULL).
 * The LCP is differentiated implicitly on its active set; contact points,
 * normals and depths are differentiated through the colliders.  The batch
 * state is left at the end of the tape. */
ODE_API int dBatchStepBackward(dBatchID b, int nsteps, const float *grad_out, float *grad_in,
                               float *grad_mu, float *grad_bounce);
/* Single-world form of the adjoint for the ODE API path (no counterpart in
 * the reference; the batch form is dBatchStepBackward).  With tape recording
 * on, each dWorldQuickStep of `w` keeps its pre-step state, its contacts
 * and its joints with their impulses; dWorldQuickStepBackward then returns
 * the derivative
 * of the last such step: grad_out = dL/d(state after), grad_in receives
 * dL/d(state before), 13 dReals per body in creation order (pos 3, q 4,
 * lvel 3, avel 3).  The LCP is differentiated implicitly at the step's
 * impulses on their active set, which is the derivative of the step when
 * its solve converged (raise dWorldSetQuickStepNumIterations until it does).
 * Differentiable steps: contact joints, ball joints, hinges without limits
 * and hinge2 joints (not a bouncy or motor-loaded one-sided limit), one
 * surface policy shared by every contact (dContactApprox1 friction or
 * mu = 0, optional dContactBounce), no forces besides gravity, contact geoms
 * in the world's space.  A step outside these conditions is still taken;
 * dWorldQuickStepBackward then reports why and returns 0, as it does when
 * nothing was recorded or the world changed since.  Returns 1 on success. */
ODE_API int dWorldSetTape(dWorldID w, int on);
ODE_API int dWorldQuickStepBackward(dWorldID w, const dReal *grad_out, dReal *grad_in);

/* library information */
ODE_API const char *odemiVersion(void);
ODE_API int odemiDeviceCount(void);

#ifdef __cplusplus
}
#endif
#endif /* ODEMI_H */
