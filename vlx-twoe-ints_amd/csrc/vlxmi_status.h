/* Status codes returned by every vlxmi C-ABI entry point. */
#ifndef VLXMI_STATUS_H
#define VLXMI_STATUS_H

/* 0 = success; negative = argument/layout error detected before any launch;
 * positive = the hipError_t reported by the HIP runtime after the launch. */
#define VLXMI_OK 0
#define VLXMI_ERR_ARG (-1)
#define VLXMI_ERR_LAYOUT (-2)
#define VLXMI_ERR_SIZE (-3)

#endif
