/*
 * cq_revolve.h -- binomial checkpointing schedule (libcq_host.so).
 * Replaces revolve()/numforw() of sgpstd-qmodeling/src.c:189-470 and the
 * action enum of sgpstd-qmodeling/revolve.h:6 (same numbering).
 */
#ifndef CQ_REVOLVE_H
#define CQ_REVOLVE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CQH_RV_ADVANCE = 0,
    CQH_RV_TAKESHOT = 1,
    CQH_RV_RESTORE = 2,
    CQH_RV_FIRSTURN = 3,
    CQH_RV_YOUTURN = 4,
    CQH_RV_TERMINATE = 5,
    CQH_RV_ERROR = 6
};

typedef struct {
    int check, capo, fine, snaps, info;
    int turn, oldsnaps, oldfine;
    int ch[64];
    int advances, takeshots, commands;
} cqh_revolve;

/* numforw (src.c:189-219): forward steps the schedule will run, or -1. */
int cqh_numforw(int steps, int snaps);
/* State of src.c's caller before the loop (sgpstd-qrtm/acps4.cu:303-304):
 * check=-1, capo=0, fine=steps, info=3. */
void cqh_revolve_init(cqh_revolve *r, int steps, int snaps);
/* One revolve() call; returns a CQH_RV_* action. */
int cqh_revolve_next(cqh_revolve *r);

#ifdef __cplusplus
}
#endif
#endif
