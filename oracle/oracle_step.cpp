/*
 * ORACLE -- test infrastructure only (see oracle_math.h).
 * One b3GpuRigidBodyPipeline::stepSimulation (b3GpuRigidBodyPipeline.cpp:221-463):
 * world AABBs, broadphase, narrowphase, contact solve, integration.
 */
#include "oracle.h"

static int g_threads = 1;
extern "C" void b3o_set_threads(int n) { g_threads = n > 0 ? n : 1; }
extern "C" int b3o_threads(void) { return g_threads; }

extern "C" int b3o_step(b3oWorld* w, float dt)
{
	b3o_compute_world_aabbs(w->numBodies, w->bodies, w->collidables, w->localAabbs, w->worldAabbs);
	int np;
	if (w->broadphase == 0)
		np = b3o_grid_pairs(w->worldAabbs, w->smallMap, w->nSmall, w->largeMap, w->nLarge, w->cellSize, w->gridDims,
							w->maxBodiesPerCell, w->pairs, w->maxPairs, 0);
	else if (w->broadphase == 1)
		np = b3o_sap_pairs(w->worldAabbs, w->smallMap, w->nSmall, w->largeMap, w->nLarge, -1, w->pairs, w->maxPairs, 0, 0);
	else
		np = b3o_plbvh_pairs(w->worldAabbs, w->smallMap, w->nSmall, w->largeMap, w->nLarge, w->pairs, w->maxPairs, 0, 0);
	w->numPairs = np;
	int nc = 0;
	if (np)
	{
		nc = b3o_narrowphase_ex(w->pairs, np, w->bodies, w->collidables, w->convexes, w->vertices, w->uniqueEdges,
								w->faces, w->indices, w->contacts, w->maxContacts, 0, 0, &w->compound, 0, 0);
		if (nc > w->maxContacts) nc = w->maxContacts;
	}
	w->numContacts = nc;
	/* joints before the contacts (b3GpuRigidBodyPipeline.cpp:361-372) */
	if (w->numConstraints > 0 && w->constraints)
		b3o_solve_joints(w->numBodies, w->bodies, w->inertias, w->constraints, w->numConstraints, 1.f / 60.f, 4);
	if (nc)
		b3o_solve_contacts(w->numBodies, w->bodies, w->inertias, nc, w->contacts, &w->solver, w->staticIdx, 0, 0, 0, 0, 0,
						   0, w->residuals);
	b3o_integrate(w->numBodies, w->bodies, dt, w->solver.m_angularDamping, w->gravity);
	return 0;
}
