/*
 * dna_tables.h — in-memory layout of the Geant4-DNA cross-section tables.
 *
 * The file cross_sections.bin is exactly this struct (little-endian f32, no padding),
 * produced from the reference's public/cross_sections.wgsl:9-72 by
 * webgpu-dna_amd/data/extract_tables.py.  Array names follow the WGSL constants.
 */
#ifndef DNA_TABLES_H
#define DNA_TABLES_H

#define DNA_XN    100   /* log energy grid, 8 eV .. 30 keV        */
#define DNA_N_AE  25    /* Champion angular CDF energies          */
#define DNA_N_AC  25    /* Champion angular CDF bins              */
#define DNA_N_WE  58    /* Born differential CDF primary energies */
#define DNA_N_WR  250   /* Born CDF breakpoints per energy        */
#define DNA_N_VIB 38    /* Sanche vibrational energy grid         */

typedef struct dna_tables {
    float XE[DNA_XN], XI[DNA_XN], XC[DNA_XN], XL[DNA_XN];
    float XSF[5][DNA_XN];                 /* ionisation shell fractions      */
    float XEF[5][DNA_XN];                 /* excitation level fractions      */
    float XAE[DNA_N_AE], XAC[DNA_N_AE * DNA_N_AC];
    float XWE[DNA_N_WE];
    float XWC[5][DNA_N_WE * DNA_N_WR];    /* Born CDF values                 */
    float XWT[5][DNA_N_WE * DNA_N_WR];    /* Born energy transfer at the CDF */
    float XVE[DNA_N_VIB], XVS[DNA_N_VIB], XVMF[DNA_N_VIB * 9], VIB_LEV[9];
    float LOG_XE0, INV_LOG_XE_STEP, LOG_XAE0, INV_LOG_XAE_STEP, LOG_XWE0, INV_LOG_XWE_STEP;
} dna_tables;

#define DNA_TABLES_BYTES 590164u

#endif
