/* rcabeam C ABI, implemented for AMD Instinct MI355X (gfx950) in HIP.
 *
 * Drop-in for the reference header include/rcabeam.h: the same 15 symbols,
 * enum values, argument order and argument meaning.  Each declaration cites the
 * reference line it replaces.  Conventions (as in the reference):
 *   - every pointer is a device pointer, every launch goes to the default
 *     (null) stream and the call returns once the work is enqueued;
 *   - the return value reports argument errors and launch errors only;
 *   - the caller owns all memory; no entry point allocates.
 * Differences that do not change results are stated next to the entry point
 * they concern (the geometry workspace is accepted but no longer used: time of
 * flight is computed on the fly in float64 per voxel/element).
 */
#pragma once

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* reference include/rcabeam.h:9-13 */
typedef enum rcabeam_status {
    RCABEAM_SUCCESS = 0,
    RCABEAM_ERROR_CUDA = 1, /* device runtime (HIP) error */
    RCABEAM_ERROR_ARGUMENT = 2,
} rcabeam_status;

/* reference include/rcabeam.h:15 */
const char* rcabeam_status_string(rcabeam_status status);

/* Angle sum of pre-beamformed per-angle IQ: iq complex64 (n_voxels, n_angles,
 * n_frames) -> out complex64 (n_voxels, n_frames).  reference include/rcabeam.h:17-23 */
rcabeam_status rcabeam_opw_device(const void* iq, void* out, size_t n_voxels, size_t n_angles, size_t n_frames);

/* |mean_t (sum RC span) * conj(sum CR span)| per voxel.  reference include/rcabeam.h:25-35 */
rcabeam_status rcabeam_xdoppler_pd_device(const void* iq, float* out, size_t n_voxels, size_t n_angles,
                                          size_t n_frames, size_t rc_start, size_t rc_count, size_t cr_start,
                                          size_t cr_count);

/* mean_t |(sum sgnsqrt RC span) * (sum sgnsqrt CR span)|^2.  reference include/rcabeam.h:37-47 */
rcabeam_status rcabeam_rc_fmas_pd_device(const void* iq, float* out, size_t n_voxels, size_t n_angles,
                                         size_t n_frames, size_t rc_start, size_t rc_count, size_t cr_start,
                                         size_t cr_count);

/* One frame, one aperture: host-layout channels complex64 (n_samples,
 * n_channels, n_angles) -> per-angle DAS out complex64 (n_voxels, n_angles).
 * config 0 = RC, 1 = CR; f_number <= 0 is the full aperture.
 * reference include/rcabeam.h:49-65 */
rcabeam_status rcabeam_delay_rca_channels_device(const void* iq_ch, const float* scan_coords_m,
                                                 const float* elements_m, const float* angles_rad,
                                                 const float* t_start_s, void* out, size_t n_samples,
                                                 size_t n_channels, size_t n_angles, size_t n_voxels, int config,
                                                 float sound_speed_m_s, float sampling_freq_hz,
                                                 float demod_freq_hz, float f_number);

/* As above, also the angle-compounded per-channel sums channel_sum complex64
 * (n_voxels, n_channels) and the active-channel count n_active (n_voxels).
 * reference include/rcabeam.h:67-85 */
rcabeam_status rcabeam_delay_rca_channel_data_device(const void* iq_ch, const float* scan_coords_m,
                                                     const float* elements_m, const float* angles_rad,
                                                     const float* t_start_s, void* channel_sum, void* angle_sum,
                                                     float* n_active, size_t n_samples, size_t n_channels,
                                                     size_t n_angles, size_t n_voxels, int config,
                                                     float sound_speed_m_s, float sampling_freq_hz,
                                                     float demod_freq_hz, float f_number);

/* One-frame fused powers from both apertures (host layout per aperture).
 * reference include/rcabeam.h:87-104 (OPW), 106-123 (XDoppler), 125-142
 * (RC-FMAS), 144-163 (all three), 165-182 (DMAS-CCF-ACF). */
rcabeam_status rcabeam_opw_pd_from_channels_device(const void* iq_rc, const void* iq_cr, const float* scan_coords_m,
                                                   const float* x_elements_m, const float* y_elements_m,
                                                   const float* angles_rad, const float* t_start_s, float* out,
                                                   size_t n_samples, size_t n_channels, size_t n_angles,
                                                   size_t n_voxels, float sound_speed_m_s, float sampling_freq_hz,
                                                   float demod_freq_hz, float f_number);

rcabeam_status rcabeam_xdoppler_pd_from_channels_device(const void* iq_rc, const void* iq_cr,
                                                        const float* scan_coords_m, const float* x_elements_m,
                                                        const float* y_elements_m, const float* angles_rad,
                                                        const float* t_start_s, float* out, size_t n_samples,
                                                        size_t n_channels, size_t n_angles, size_t n_voxels,
                                                        float sound_speed_m_s, float sampling_freq_hz,
                                                        float demod_freq_hz, float f_number);

rcabeam_status rcabeam_rc_fmas_pd_from_channels_device(const void* iq_rc, const void* iq_cr,
                                                       const float* scan_coords_m, const float* x_elements_m,
                                                       const float* y_elements_m, const float* angles_rad,
                                                       const float* t_start_s, float* out, size_t n_samples,
                                                       size_t n_channels, size_t n_angles, size_t n_voxels,
                                                       float sound_speed_m_s, float sampling_freq_hz,
                                                       float demod_freq_hz, float f_number);

rcabeam_status rcabeam_fast_pd_from_channels_device(const void* iq_rc, const void* iq_cr, const float* scan_coords_m,
                                                    const float* x_elements_m, const float* y_elements_m,
                                                    const float* angles_rad, const float* t_start_s, float* opw_out,
                                                    float* xdoppler_out, float* rc_fmas_out, size_t n_samples,
                                                    size_t n_channels, size_t n_angles, size_t n_voxels,
                                                    float sound_speed_m_s, float sampling_freq_hz,
                                                    float demod_freq_hz, float f_number);

rcabeam_status rcabeam_dmas_ccf_acf_from_channels_device(const void* iq_rc, const void* iq_cr,
                                                         const float* scan_coords_m, const float* x_elements_m,
                                                         const float* y_elements_m, const float* angles_rad,
                                                         const float* t_start_s, float* out, size_t n_samples,
                                                         size_t n_channels, size_t n_angles, size_t n_voxels,
                                                         float sound_speed_m_s, float sampling_freq_hz,
                                                         float demod_freq_hz, float f_number);

/* Convert complex64 (samples, channels, angles, frames) to the packed
 * frame-inner layout (channels, angles, samples, frames).
 * reference include/rcabeam.h:184-185 */
rcabeam_status rcabeam_pack_ensemble_device(const void* source, void* target, size_t ns, size_t nc, size_t na,
                                            size_t nt);

/* Frame-batched reconstruction from packed complex64 RC/CR channels.
 * method: 0 OPW, 1 XDoppler, 2 RC-FMAS, 3 DMAS-CCF-ACF, 4 St-SW, 5 separate
 * aperture IQ.  pd: 2*nv floats, the result in the even entries (methods 1 and
 * 4 leave the imaginary part of the mean cross-product in the odd entries,
 * other methods write zero there).  signal: optional complex64 (nv, nt) for
 * methods 0-3 (method 3: real DMAS response, imaginary zero); method 4 needs a
 * null signal; method 5 needs complex64 (nv, 2, nt), RC then CR, and sets pd
 * to zero.  St-SW needs weights[nv], complex64 workspace[nv*k*k*nt] and
 * 2 <= k <= min(4, na); normalise the weights by the full-volume maximum
 * before multiplying the power.
 * geometry_workspace: must be non-null (reference contract); it is neither
 * read nor written here.  Results do not depend on alignment or on the
 * parity of nt.  reference include/rcabeam.h:187-207 */
rcabeam_status rcabeam_ensemble_device(const void* rc, const void* cr, const float* scan, const float* xe,
                                       const float* ye, const float* angles, const float* starts, float* pd,
                                       void* signal, float* weights, void* workspace, void* geometry_workspace,
                                       size_t ns, size_t nc, size_t na, size_t nt, size_t nv, float c, float fs,
                                       float fd, float fn, int method, int k);

/* Pack complex64 channels to complex FP16 (two IEEE binary16 components), same
 * layout conversion as rcabeam_pack_ensemble_device.  range_error points to one
 * device int the caller zeroes; non-finite components or components beyond
 * +/-65504 set it to one (underflow and rounding are not errors).  Source and
 * target must not alias.  Source 8-byte aligned; target and range_error
 * 4-byte aligned.  reference include/rcabeam.h:209-220 */
rcabeam_status rcabeam_pack_ensemble_fp16_device(const void* source, void* target, int* range_error, size_t ns,
                                                 size_t nc, size_t na, size_t nt);

/* OPW from packed FP16 RC/CR (4 bytes per complex value) with float32
 * interpolation and accumulation; pd and the optional complex64 signal as in
 * rcabeam_ensemble_device.  Packed inputs need 4-byte alignment.
 * reference include/rcabeam.h:222-235 */
rcabeam_status rcabeam_opw_fp16_device(const void* rc, const void* cr, const float* scan, const float* xe,
                                       const float* ye, const float* angles, const float* starts, float* pd,
                                       void* signal, void* geometry_workspace, size_t ns, size_t nc, size_t na,
                                       size_t nt, size_t nv, float c, float fs, float fd, float fn);

#ifdef __cplusplus
}
#endif
