/* Internal interface between the drop-in host layer and its helpers. */
#ifndef LPM_DROPIN_H
#define LPM_DROPIN_H

#include "lpmx.h"

/* Abort like the reference's fatal paths: message, then exit(1). */
void lpm_dropin_fatal(const char *where, int rc);

/* Transfers between the reference's host globals and the device fields. */
void lpm_dropin_push(int field);
void lpm_dropin_pull(int field);
/* resident mode: host code about to read (write 0) or change (1) the array at p, or
 * reading many arrays (sync_host): the stale host copies come back first */
void lpm_dropin_host_access(const void *p, int write);
void lpm_dropin_sync_host(void);

/* Hand slipSysDefine3D's Schmid tensors and the cp_* parameters to the device. */
void lpm_dropin_slip_systems(void);

#endif
