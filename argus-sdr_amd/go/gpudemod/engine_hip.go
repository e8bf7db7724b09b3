//go:build hip

// Package gpudemod, hip build tag: the Engine and BatchRunner streaming
// surface of internal/demod/gpudemod over libargus_hip.so (include/argus_gpu.h).
// Drop-in beside the cufft files: build with -tags hip and
//   CGO_CFLAGS="-I<argus>/include" CGO_LDFLAGS="-L<argus>/argus-sdr_amd -largus_hip -Wl,-rpath,<argus>/argus-sdr_amd".
package gpudemod

/*
#cgo LDFLAGS: -largus_hip
#include <stdlib.h>
#include "argus_gpu.h"
*/
import "C"

import (
	"errors"
	"math"
	"unsafe"

	"sdr-wideband-suite/internal/dsp"
)

type DemodType int

const (
	DemodNFM DemodType = iota
	DemodWFM
	DemodAM
	DemodUSB
	DemodLSB
	DemodCW
)

// Engine keeps one device buffer per stage; a call uploads its input once,
// chains the stages on the device and downloads the result once (the cufft
// Engine round-trips through the host after every stage, gpudemod.go:184-331).
type Engine struct {
	maxSamples       int
	sampleRate       int
	phase            float64
	bfoPhase         float64
	firTaps          []float32
	lastShiftUsedGPU bool
	lastDemodUsedGPU bool
	dIn, dShifted    unsafe.Pointer
	dFiltered, dDec  unsafe.Pointer
	dAudio           unsafe.Pointer
	runner           *C.gpud_runner // streaming batch runner (StreamingExtractGPU)
}

func Available() bool {
	var n C.int
	return C.gpud_device_count(&n) == 0 && n > 0
}

func New(maxSamples int, sampleRate int) (*Engine, error) {
	if maxSamples <= 0 {
		return nil, errors.New("invalid maxSamples")
	}
	if sampleRate <= 0 {
		return nil, errors.New("invalid sampleRate")
	}
	if !Available() {
		return nil, errors.New("hip device not available")
	}
	e := &Engine{maxSamples: maxSamples, sampleRate: sampleRate}
	iq := C.size_t(maxSamples) * C.size_t(unsafe.Sizeof(complex64(0)))
	for _, p := range []*unsafe.Pointer{&e.dIn, &e.dShifted, &e.dFiltered, &e.dDec} {
		if C.gpud_malloc(p, iq) != 0 {
			e.Close()
			return nil, ErrUnavailable
		}
	}
	if C.gpud_malloc(&e.dAudio, C.size_t(maxSamples)*4) != 0 {
		e.Close()
		return nil, ErrUnavailable
	}
	if C.gpud_runner_create(C.int(maxSamples), C.int(sampleRate), &e.runner) != 0 {
		e.Close()
		return nil, ErrUnavailable
	}
	return e, nil
}

func (e *Engine) SetFIR(taps []float32) {
	if len(taps) == 0 {
		e.firTaps = nil
		return
	}
	if len(taps) > 256 {
		taps = taps[:256]
	}
	e.firTaps = append(e.firTaps[:0], taps...)
	_ = C.gpud_upload_fir_taps_hip((*C.float)(unsafe.Pointer(&e.firTaps[0])), C.int(len(e.firTaps)))
}

func (e *Engine) LastShiftUsedGPU() bool { return e != nil && e.lastShiftUsedGPU }
func (e *Engine) LastDemodUsedGPU() bool { return e != nil && e.lastDemodUsedGPU }

func (e *Engine) upload(iq []complex64) bool {
	return C.gpud_memcpy_h2d(e.dIn, unsafe.Pointer(&iq[0]), C.size_t(len(iq))*8) == 0
}

// chain runs shift (phase_start, phase_inc), FIR with the uploaded taps and
// decimation on the device; returns the decimated length: ceil(n/decim) as
// dsp.Decimate keeps (fir.go:84-95), or floor(n/decim) as the cufft engine's
// decimate stage does (gpudemod.go:258).
func (e *Engine) chain(n int, phaseInc, phaseStart float64, numTaps, decim int, ceil bool) (int, bool) {
	in, sh, fi, de := (*C.gpud_float2)(e.dIn), (*C.gpud_float2)(e.dShifted), (*C.gpud_float2)(e.dFiltered), (*C.gpud_float2)(e.dDec)
	if C.gpud_launch_freq_shift_hip(in, sh, C.int(n), C.double(phaseInc), C.double(phaseStart)) != 0 {
		return 0, false
	}
	if C.gpud_launch_fir_hip(sh, fi, C.int(n), C.int(numTaps)) != 0 {
		return 0, false
	}
	nOut := n / decim
	if ceil {
		nOut = (n + decim - 1) / decim
	}
	if nOut == 0 {
		return 0, true
	}
	if C.gpud_launch_decimate_hip(fi, de, C.int(nOut), C.int(decim)) != 0 {
		return 0, false
	}
	return nOut, true
}

// ShiftFilterDecimate (gpudemod.go:333-357): FreqShift from phase 0 with the
// increment applied before the first sample, 101-tap LowpassFIR at
// max(bw/2, 200) Hz (clamped below Nyquist), keep every round(fs/outRate)-th.
func (e *Engine) ShiftFilterDecimate(iq []complex64, offsetHz float64, bw float64, outRate int) ([]complex64, int, error) {
	if e == nil {
		return nil, 0, errors.New("nil HIP demod engine")
	}
	if len(iq) == 0 {
		return nil, 0, nil
	}
	if outRate <= 0 {
		return nil, 0, errors.New("invalid output sample rate")
	}
	if len(iq) > e.maxSamples {
		return nil, 0, errors.New("sample count exceeds engine capacity")
	}
	cutoff := math.Max(bw/2, 200)
	if cutoff > float64(e.sampleRate)/2-1 {
		cutoff = float64(e.sampleRate)/2 - 1
	}
	t64 := dsp.LowpassFIR(cutoff, e.sampleRate, 101)
	taps := make([]float32, len(t64))
	for i, v := range t64 {
		taps[i] = float32(v)
	}
	e.SetFIR(taps)
	decim := int(math.Round(float64(e.sampleRate) / float64(outRate)))
	if decim < 1 {
		decim = 1
	}
	inc := -2 * math.Pi * offsetHz / float64(e.sampleRate)
	if !e.upload(iq) {
		return nil, 0, ErrUnavailable
	}
	nOut, ok := e.chain(len(iq), inc, inc, len(taps), decim, true)
	if !ok {
		return nil, 0, ErrUnavailable
	}
	out := make([]complex64, nOut)
	if nOut > 0 && C.gpud_memcpy_d2h(unsafe.Pointer(&out[0]), e.dDec, C.size_t(nOut)*8) != 0 {
		return nil, 0, ErrUnavailable
	}
	e.lastShiftUsedGPU = true
	return out, e.sampleRate / decim, nil
}

func (e *Engine) DemodFused(iq []complex64, offsetHz float64, bw float64, mode DemodType) ([]float32, int, error) {
	return e.Demod(iq, offsetHz, bw, mode)
}

// Demod (gpudemod.go:364-461): carried shift phase, the engine's FIR (or a
// 101-tap design at max(bw/2, 200) Hz), decimation to 48 kHz (192 kHz WFM)
// and the discriminator / envelope / product detector with a carried BFO.
func (e *Engine) Demod(iq []complex64, offsetHz float64, bw float64, mode DemodType) ([]float32, int, error) {
	if e == nil {
		return nil, 0, errors.New("nil HIP demod engine")
	}
	if len(iq) == 0 {
		return nil, 0, nil
	}
	if len(iq) > e.maxSamples {
		return nil, 0, errors.New("sample count exceeds engine capacity")
	}
	outRate := 48000
	switch mode {
	case DemodNFM, DemodAM, DemodUSB, DemodLSB, DemodCW:
	case DemodWFM:
		outRate = 192000
	default:
		return nil, 0, errors.New("unsupported demod type")
	}
	if len(e.firTaps) == 0 {
		t64 := dsp.LowpassFIR(math.Max(bw/2, 200), e.sampleRate, 101)
		taps := make([]float32, len(t64))
		for i, v := range t64 {
			taps[i] = float32(v)
		}
		e.SetFIR(taps)
	}
	decim := int(math.Round(float64(e.sampleRate) / float64(outRate)))
	if decim < 1 {
		decim = 1
	}
	inc := -2 * math.Pi * offsetHz / float64(e.sampleRate)
	if !e.upload(iq) {
		return nil, 0, ErrUnavailable
	}
	nDec, ok := e.chain(len(iq), inc, e.phase, len(e.firTaps), decim, false)
	if !ok {
		return nil, 0, ErrUnavailable
	}
	e.phase += inc * float64(len(iq))
	e.lastShiftUsedGPU = true
	rate := e.sampleRate / decim
	de, au := (*C.gpud_float2)(e.dDec), (*C.float)(e.dAudio)
	nAudio := nDec
	var rc C.int
	switch mode {
	case DemodNFM, DemodWFM:
		if nDec < 2 {
			return nil, rate, nil
		}
		rc = C.gpud_launch_fm_discrim_hip(de, au, C.int(nDec))
		nAudio = nDec - 1
	case DemodAM:
		rc = C.gpud_launch_am_envelope_hip(de, au, C.int(nDec))
	default:
		bfo := 700.0
		if mode == DemodLSB {
			bfo = -700.0
		}
		bInc := 2 * math.Pi * bfo / float64(rate)
		rc = C.gpud_launch_ssb_product_hip(de, au, C.int(nDec), C.double(bInc), C.double(e.bfoPhase))
		e.bfoPhase += bInc * float64(nDec)
	}
	if rc != 0 || C.gpud_device_sync() != 0 {
		return nil, 0, ErrUnavailable
	}
	out := make([]float32, nAudio)
	if nAudio > 0 && C.gpud_memcpy_d2h(unsafe.Pointer(&out[0]), e.dAudio, C.size_t(nAudio)*4) != 0 {
		return nil, 0, ErrUnavailable
	}
	e.lastDemodUsedGPU = true
	return out, rate, nil
}

func (e *Engine) Close() {
	if e == nil {
		return
	}
	for _, p := range []*unsafe.Pointer{&e.dIn, &e.dShifted, &e.dFiltered, &e.dDec, &e.dAudio} {
		if *p != nil {
			_ = C.gpud_free(*p)
			*p = nil
		}
	}
	if e.runner != nil {
		_ = C.gpud_runner_destroy(e.runner)
		e.runner = nil
	}
}
