/*
 * rca_sim.h — C ABI of the MI355X (gfx950) SIMUS-style RCA simulator, the
 * upstream generator of the beamformer's I/Q (MATLAB/lib/simulator/
 * SimUS_CUDARCA.cu and .m, with the receive-aberration variant _ab).
 *
 * Plain C: device pointers, sizes and scalars; complex arrays are interleaved
 * float (re, im) pairs ("complex64").  Every call is asynchronous on the
 * given stream unless stated.  Errors return a non-zero rca_status (see
 * rca_das.h) and leave a message in rca_last_error().
 *
 * Array layouts are the reference's column-major MATLAB arrays, written here
 * as C arrays with the last MATLAB index first:
 *   r_ele   float [nSub][5][nEle]    ElementPos [nEle, 5, nSub]: x, y, z (m),
 *                                    azimuth, elevation (rad) per sub-element
 *   r_scat  float [nTx][3][nScat]    scatterer positions of each tilt (m)
 *   k, spectra  complex64 [Nf]       wavenumber (rad/m + j Np/m) and spectra
 *   TX      complex64 [nTx][nScat][Nf]
 *   RX      complex64 [nScat][nEle][Nf]
 *   Delay, Apod  float [nTx][nEle]   transmit laws (Delay minus the receive
 *                                    start time, as the wrapper passes it)
 */
#ifndef RCA_SIM_H
#define RCA_SIM_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/*
 * Transmit field, reference kernel ABI: TXfield (SimUS_CUDARCA.cu:209-228,
 * launched by feval at SimUS_CUDARCA.m:314-329), same arguments and meaning,
 * plus bsc and the stream.  TX[t][s][f] = sum over elements and sub-elements
 * of Green x directivities x lens x conj(PULSE IR) x Apod exp(j w Delay).
 * bsc (float [nTx][nScat], may be NULL) applies the wrapper's next two steps
 * in the same pass: TX *= bsc, NaN -> 0 (.m:331-335).
 */
int rca_sim_txfield(void* TX, const float* r_scat, const float* r_ele, const void* kWaveNumber,
                    const void* PULSE, const void* IR, const float* Delay, const float* Apod,
                    float width, float height, float rfocus, float baffle, float lensCorrection,
                    float SpeedOfSound, int NscatPerEvent, int Nele, int Nf, int Ntx, int Nsubele,
                    const float* bsc, hipStream_t stream);

/*
 * Receive field, reference kernel ABI: RXfield (SimUS_CUDARCA.cu:301-316,
 * launched at SimUS_CUDARCA.m:340-355), plus the stream.  r_scat is one
 * tilt's [3][nScat].  RX[s][e][f] = sum over sub-elements of Green x
 * directivities x lens x IR.
 */
int rca_sim_rxfield(void* RX, const float* r_scat, const float* r_ele, const void* kWaveNumber,
                    const void* IR, float width, float height, float rfocus, float baffle,
                    float lensCorrection, float SpeedOfSound, int Nscat, int Nele, int Nf,
                    int Nsubele, hipStream_t stream);

/*
 * Fused receive (no single reference counterpart): for every tilt t of a
 * frame, the receive field of its scatterers, the receive apodisation, the
 * sum over scatterers against TX and the input filter, conjugated and
 * zero-padded for the inverse transform - the RXfield feval, .*Apod,
 * sum(TX.*RX,3).*IPFilt and conj of SimUS_CUDARCA.m:338-363 - without
 * materialising RX.  With ab_delay/ab_apod it also writes the aberrated
 * spectra of SimUS_CUDARCA_ab.m:365-376.
 */
typedef struct {
  const void* tx;          /* TX field complex64 [nTx][nScat][Nf] (bsc applied) */
  const float* r_scat;     /* [nTx][3][nScat] */
  const float* r_ele;      /* receive array [nSub][5][nEle] */
  const void* k;           /* complex64 [Nf] */
  const void* ir;          /* receive impulse-response spectrum complex64 [Nf] */
  const void* ipfilt;      /* input-filter spectrum complex64 [Nf] */
  const float* agl;        /* angular frequencies 2 pi f (rad/s) [Nf], for the aberration law */
  const float* apod;       /* receive apodisation [nTx][nEle] */
  const float* ab_delay;   /* receive aberration delays (s) [nEle], or NULL */
  const float* ab_apod;    /* receive aberration apodisation [nEle], or NULL */
  float width, height, rfocus, baffle, lens, speed_of_sound;
  int n_scat, n_ele, n_f, n_sub, n_tx, nfft;
  void* spec;              /* out: complex64 [nTx][nEle][nfft] */
  void* spec_ab;           /* out: same for the aberrated receive, or NULL */
} rca_sim_receive_args;

int rca_sim_receive(const rca_sim_receive_args* args, hipStream_t stream);

/*
 * Time-domain output (SimUS_CUDARCA.m:363-390): in place, the inverse FFT of
 * each of the n_batch spectra (complex64 [n_batch][nfft], overwritten), its
 * real part / nfft, trimmed or zero-padded to n_sample rows, and of each
 * `sub` samples the first two kept:
 * raw[b][2i] = RF[i*sub], raw[b][2i+1] = RF[i*sub + 1], i < n_sample/sub.
 * raw is float [n_batch][2*(n_sample/sub)] (MATLAB Raw [2*n_sample/sub,
 * nEle, nEvent]).  Uses hipFFT (plans cached per device and shape).
 */
int rca_sim_spectra_to_raw(void* spec, int n_batch, int nfft, int n_sample, int sub, float* raw,
                           hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* RCA_SIM_H */
