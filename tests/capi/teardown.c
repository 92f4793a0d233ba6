/* Host-only checks of the ODE API object model (no step, no device work),
 * built with AddressSanitizer against a host-only build of the library:
 *
 *  1. dBodyDestroy detaches the body's geoms the way dGeomSetBody(g, 0) does
 *     (ode.cpp:296-309, collision_kernel.cpp:530-547): the geom keeps the
 *     body's last pose as its own.
 *  2. dWorldDestroy before dJointGroupDestroy and dSpaceDestroy (the
 *     teardown order of demos/raycar/ode.cpp:26-29) leaves no geom, joint or
 *     space pointing at freed memory (ode.cpp:1585-1612).
 *
 * Prints "ok" and exits 0 when every check holds. */
#include <math.h>
#include <stdio.h>

#include "odemi.h"

static int fails = 0;
#define CHECK(c, msg)                                   \
  do {                                                  \
    if (!(c)) { printf("FAIL: %s\n", msg); fails++; }   \
  } while (0)

int main(void) {
  dInitODE2(0);
  /* 1: body destroyed, geom survives at the body's last pose */
  {
    dWorldID w = dWorldCreate();
    dSpaceID s = dHashSpaceCreate(0);
    dBodyID b = dBodyCreate(w);
    dBodySetPosition(b, 1.5f, -2.0f, 3.25f);
    dQuaternion q = {0.8f, 0.6f, 0.0f, 0.0f};
    dBodySetQuaternion(b, q);
    dGeomID g = dCreateBox(s, 1, 1, 1);
    dGeomSetBody(g, b);
    dReal R0[12];
    const dReal *Rb = dBodyGetRotation(b);
    for (int k = 0; k < 12; k++) R0[k] = Rb[k];
    dBodyDestroy(b);
    CHECK(dGeomGetBody(g) == 0, "geom still attached after dBodyDestroy");
    const dReal *p = dGeomGetPosition(g);
    CHECK(p[0] == 1.5f && p[1] == -2.0f && p[2] == 3.25f, "geom lost the body's position");
    const dReal *R = dGeomGetRotation(g);
    int same = 1;
    for (int k = 0; k < 12; k++)
      if (k % 4 != 3 && R[k] != R0[k]) same = 0;
    CHECK(same, "geom lost the body's rotation");
    dSpaceDestroy(s);
    dWorldDestroy(w);
  }
  /* 2: world first, then the joint group, then the space */
  {
    dWorldID w = dWorldCreate();
    dSpaceID s = dHashSpaceCreate(0);
    dJointGroupID cg = dJointGroupCreate(0);
    dGeomID plane = dCreatePlane(s, 0, 0, 1, 0);
    (void)plane;
    dBodyID bs[4];
    dGeomID gs[4];
    dMass m;
    for (int i = 0; i < 4; i++) {
      bs[i] = dBodyCreate(w);
      dBodySetPosition(bs[i], (dReal)i, 0, 0.5f + i);
      dMassSetBox(&m, 1, 1, 1, 1);
      dBodySetMass(bs[i], &m);
      gs[i] = dCreateBox(s, 1, 1, 1);
      dGeomSetBody(gs[i], bs[i]);
    }
    dJointID ball = dJointCreateBall(w, 0);
    dJointAttach(ball, bs[0], bs[1]);
    dJointSetBallAnchor(ball, 0.5f, 0, 1);
    dContact c;
    for (int k = 0; k < (int)sizeof(c); k++) ((char *)&c)[k] = 0;
    c.surface.mode = dContactApprox1;
    c.surface.mu = 0.5f;
    c.geom.normal[2] = 1;
    c.geom.g1 = gs[2];
    c.geom.g2 = gs[3];
    dJointID cj = dJointCreateContact(w, cg, &c);
    dJointAttach(cj, bs[2], bs[3]);
    dWorldDestroy(w);
    for (int i = 0; i < 4; i++) {
      CHECK(dGeomGetBody(gs[i]) == 0, "geom still attached after dWorldDestroy");
      const dReal *p = dGeomGetPosition(gs[i]);
      CHECK(p[0] == (dReal)i && p[2] == 0.5f + i, "geom lost its body's pose in dWorldDestroy");
    }
    dJointGroupDestroy(cg);
    dSpaceDestroy(s);
  }
  dCloseODE();
  if (fails) return 1;
  printf("ok\n");
  return 0;
}
