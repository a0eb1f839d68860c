/*
 * cq_revolve.c -- binomial checkpointing schedule for RTM source-wavefield
 * reconstruction (Griewank & Walther, ACM TOMS 26(1), 2000, Algorithm 799
 * "revolve").
 *
 * Replaces revolve()/numforw() of sgpstd-qmodeling/src.c:189-470 (identical
 * copy in sgpstd-qrtm/src.c).  The reference keeps its state in function
 * statics; here it lives in a caller-owned cqh_revolve so several shots can be
 * scheduled side by side (batched RTM).  The action sequence, the checkpoint
 * slot numbers and the info codes are the reference's, including its limits
 * checkup = repsup = 64.
 */
#include <stdio.h>
#include <string.h>

#include "cq_revolve.h"

#define RV_CHECKUP 64
#define RV_REPSUP 64

int cqh_numforw(int steps, int snaps)
{
    if (snaps < 1 || snaps > RV_CHECKUP) return -1;
    int reps = 0, range = 1;
    while (range < steps) {
        reps += 1;
        range = range * (reps + snaps) / reps;
    }
    if (reps > RV_REPSUP) return -1;
    return reps * steps - range * reps / (snaps + 1);
}

void cqh_revolve_init(cqh_revolve *r, int steps, int snaps)
{
    memset(r, 0, sizeof(*r));
    r->check = -1;
    r->capo = 0;
    r->fine = steps;
    r->snaps = snaps;
    r->info = 3;
}

/* One call of the schedule: returns the next action and updates capo, fine
 * and check exactly as revolve() does. */
int cqh_revolve_next(cqh_revolve *r)
{
    r->commands += 1;
    if (r->check < -1 || r->capo > r->fine) return CQH_RV_ERROR;
    if (r->check == -1 && r->capo < r->fine) {
        r->turn = 0;
        r->ch[0] = r->capo - 1;
    }
    const int span = r->fine - r->capo;
    if (span == 0) {
        /* nothing left in this range: step back to the previous checkpoint */
        if (r->check == -1 || r->capo == r->ch[0]) {
            r->check -= 1;
            return CQH_RV_TERMINATE;
        }
        r->capo = r->ch[r->check];
        r->oldfine = r->fine;
        return CQH_RV_RESTORE;
    }
    if (span == 1) {
        /* one combined forward / reverse step */
        r->fine -= 1;
        if (r->check >= 0 && r->ch[r->check] == r->capo) r->check -= 1;
        r->oldfine = r->fine;
        if (r->turn == 0) {
            r->turn = 1;
            return CQH_RV_FIRSTURN;
        }
        return CQH_RV_YOUTURN;
    }
    if (r->check == -1 || r->ch[r->check] != r->capo) {
        /* store the current state */
        r->check += 1;
        if (r->check >= RV_CHECKUP) { r->info = 10; return CQH_RV_ERROR; }
        if (r->check + 1 > r->snaps) { r->info = 11; return CQH_RV_ERROR; }
        r->ch[r->check] = r->capo;
        if (r->check == 0) {
            r->advances = 0;
            r->takeshots = 0;
            r->commands = 1;
            r->oldsnaps = r->snaps;
            if (r->snaps > RV_CHECKUP) { r->info = 14; return CQH_RV_ERROR; }
            if (r->info > 0 && cqh_numforw(r->fine - r->capo, r->snaps) == -1) {
                r->info = 12;
                return CQH_RV_ERROR;
            }
        }
        r->takeshots += 1;
        r->oldfine = r->fine;
        return CQH_RV_TAKESHOT;
    }
    /* advance capo by the binomially optimal amount */
    if (r->oldfine < r->fine && r->snaps == r->check + 1) { r->info = 13; return CQH_RV_ERROR; }
    const int oldcapo = r->capo;
    const int ds = r->snaps - r->check;
    if (ds < 1) { r->info = 11; return CQH_RV_ERROR; }
    int reps = 0, range = 1;
    while (range < r->fine - r->capo) {
        reps += 1;
        range = range * (reps + ds) / reps;
    }
    if (reps > RV_REPSUP) { r->info = 15; return CQH_RV_ERROR; }
    if (r->snaps != r->oldsnaps && r->snaps > RV_CHECKUP) { r->info = 14; return CQH_RV_ERROR; }
    const int b1 = range * reps / (ds + reps);
    const int b2 = (ds > 1) ? b1 * ds / (ds + reps - 1) : 1;
    const int b3 = (ds == 1) ? 0 : ((ds > 2) ? b2 * (ds - 1) / (ds + reps - 2) : 1);
    const int b4 = b2 * (reps - 1) / ds;
    const int b5 = (ds < 3) ? 0 : ((ds > 3) ? b3 * (ds - 2) / reps : 1);
    if (r->fine - r->capo <= b1 + b3)
        r->capo = r->capo + b4;
    else if (r->fine - r->capo >= range - b5)
        r->capo = r->capo + b1;
    else
        r->capo = r->fine - b2 - b3;
    if (r->capo == oldcapo) r->capo = oldcapo + 1;
    r->advances += r->capo - oldcapo;
    r->oldfine = r->fine;
    return CQH_RV_ADVANCE;
}
