/* cfm_probe.h -- measurement helper beside the mobility library (libcfm_probe.so).
 *
 * No counterpart in the reference.  Issue rate of one FP32 VALU instruction form on
 * `device`, in lane-instructions per second over the whole GPU: the ceiling against
 * which the FP32 pair kernel's rate (its VALU instructions per pair x pairs per
 * second) is judged, since every device build of the mobility library disables
 * packed FP32 instructions.  form: 0 v_fmac_f32 (two-operand, VGPR sources),
 * 1 v_mul_f32 with a scalar operand, 2 v_add_f32 with a scalar operand, 3 v_mul_f32
 * with VGPR operands, 4 v_fma_f32 (three-operand) with VGPR operands.  Allocates and
 * frees its own 8 MiB buffer and stream.  Returns 0, or 1 for a bad argument, 2 for a
 * HIP error.
 */
#ifndef CFM_PROBE_H
#define CFM_PROBE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CFM_PROBE_OK 0
#define CFM_PROBE_EINVAL 1
#define CFM_PROBE_EHIP 2

int cfm_probe_valu_rate(int32_t device, int32_t form, double* lane_instr_per_s);

#ifdef __cplusplus
}
#endif

#endif /* CFM_PROBE_H */
