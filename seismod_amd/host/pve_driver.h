/*
 * pve_driver.h -- the shared driver of the drop-in programs pve_iso_pml
 * (vti 0) and pve_vti_pml (vti 1); returns the process exit status.
 */
#ifndef SM_PVE_DRIVER_H
#define SM_PVE_DRIVER_H

int sm_pve_program(int argc, char *argv[], int vti);

#endif
