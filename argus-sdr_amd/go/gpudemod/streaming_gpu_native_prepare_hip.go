//go:build hip

package gpudemod

/*
#include "argus_gpu.h"
*/
import "C"

import "unsafe"

// executeStreamingGPUNativePrepared for the hip build: every invocation of a
// chunk goes through one gpud_runner_streaming_extract call -- one fused
// shift + polyphase FIR/decimate launch for all signals, state (NCO phase,
// decimator phase, shifted history, taps) resident on the device -- instead
// of the per-signal upload/launch/download loop of the cufft build
// (streaming_gpu_native_prepare.go:17-108).  The runner designs each
// signal's taps from the same job fields getOrInitExtractState uses
// (stream_state.go:25-85), resets a signal when its config hash changes and
// prunes signals absent from the call.
func (r *BatchRunner) executeStreamingGPUNativePrepared(invocations []StreamingGPUInvocation) ([]StreamingGPUExecutionResult, error) {
	if r == nil || r.eng == nil || r.eng.runner == nil {
		return nil, ErrUnavailable
	}
	if len(invocations) == 0 {
		return nil, nil
	}
	iq := invocations[0].IQNew
	jobs := make([]C.gpud_streaming_job, len(invocations))
	outCap := 0
	for i, inv := range invocations {
		jobs[i] = C.gpud_streaming_job{
			signal_id:   C.int64_t(inv.SignalID),
			offset_hz:   C.double(inv.OffsetHz),
			bandwidth:   C.double(inv.Bandwidth),
			out_rate:    C.int32_t(inv.OutRate),
			num_taps:    C.int32_t(inv.NumTaps),
			config_hash: C.uint64_t(inv.ConfigHash),
		}
		outCap += (inv.PhaseCountIn+len(iq))/maxInt(1, inv.Decim) + 1
	}
	out := make([]complex64, maxInt(outCap, 1))
	res := make([]C.gpud_streaming_result, len(invocations))
	var iqPtr *C.gpud_float2
	if len(iq) > 0 {
		iqPtr = (*C.gpud_float2)(unsafe.Pointer(&iq[0]))
	}
	if C.gpud_runner_streaming_extract(r.eng.runner, iqPtr, C.int(len(iq)), &jobs[0], C.int(len(jobs)),
		(*C.gpud_float2)(unsafe.Pointer(&out[0])), C.int64_t(len(out)), &res[0]) != 0 {
		return nil, ErrUnavailable
	}
	results := make([]StreamingGPUExecutionResult, len(invocations))
	for i, inv := range invocations {
		rr := res[i]
		n := int(rr.n_out)
		off := int(rr.out_offset)
		outHost := r.outRingFor(inv.SignalID).next(n, &r.outPool)
		copy(outHost, out[off:off+n])
		hist := make([]complex64, int(rr.history_len))
		if len(hist) > 0 {
			var cfg C.uint64_t
			var ph C.double
			var decim, pc, taps C.int
			C.gpud_runner_get_signal_state(r.eng.runner, C.int64_t(inv.SignalID), &cfg, &ph, &decim, &pc, &taps,
				(*C.gpud_float2)(unsafe.Pointer(&hist[0])), C.int(len(hist)))
		}
		results[i] = StreamingGPUExecutionResult{
			SignalID:      inv.SignalID,
			Mode:          StreamingGPUExecCUDA,
			IQ:            outHost,
			Rate:          int(rr.rate),
			NOut:          n,
			PhaseCountOut: int(rr.phase_count),
			NCOPhaseOut:   float64(rr.nco_phase),
			HistoryOut:    hist,
			HistoryLenOut: len(hist),
		}
	}
	return results, nil
}

// The runner prunes absent signals itself on its next call; the host-side
// output rings are pruned here as in the cufft build.
func (r *BatchRunner) syncNativeStreamingStates(active map[int64]struct{}) {
	if r == nil {
		return
	}
	r.pruneOutRings(active)
}

func (r *BatchRunner) resetNativeStreamingState(signalID int64) {
	if r == nil || r.eng == nil || r.eng.runner == nil {
		return
	}
	C.gpud_runner_reset_signal_state(r.eng.runner, C.int64_t(signalID))
}

func (r *BatchRunner) resetAllNativeStreamingStates() {
	if r == nil || r.eng == nil || r.eng.runner == nil {
		return
	}
	C.gpud_runner_reset_all_signal_states(r.eng.runner)
}

func (r *BatchRunner) freeAllNativeStreamingStates() {
	if r == nil || r.eng == nil || r.eng.runner == nil {
		return
	}
	C.gpud_runner_reset_all_signal_states(r.eng.runner)
}
