"""Input decks of the reference's documented runs (docs/examples/*.sp), the
circuits and analysis cards as the docs give them, with the outputs the
reference printed for them (docs/learn/measure.md:24-36, analyses.md:25-30,
models.md:38-45, 95-102)."""

MEASURE_SP = """Measurements on an RC step response
* first-order RC, tau = 1 us; the input steps up at 1 us and down at 6 us
V1 in 0 PULSE(0 1 1u 1n 1n 5u 20u)
R1 in out 1k
C1 out 0 1n
.tran 10n 10u
.meas tran trise TRIG v(out) VAL=0.1 RISE=1 TARG v(out) VAL=0.9 RISE=1
.meas tran tdelay TRIG v(in) VAL=0.5 RISE=1 TARG v(out) VAL=0.5 RISE=1
.meas tran vat3u FIND v(out) AT=3u
.meas tran tcross WHEN v(out)=0.5 FALL=1
.meas tran vmax MAX v(out)
.meas tran vavg AVG v(out) FROM=0 TO=10u
.meas tran vrms RMS v(out) FROM=0 TO=10u
.meas tran slope DERIV v(out) AT=2u
.meas tran charge INTEG i(V1) FROM=1u TO=6u
.meas tran ratio PARAM='trise/tdelay'
.end
"""

# name -> (value, printed digits) from docs/learn/measure.md:24-36; the
# side values (targ/trig, at, from/to) follow as name.side.
MEASURE_GOLDEN = {
    "trise": (2.197201e-06, 6), "tdelay": (6.931367e-07, 6), "vat3u": (8.645984e-01, 6),
    "tcross": (6.68790e-06, 5), "vmax": (9.932657e-01, 6), "vavg": (4.982781e-01, 6),
    "vrms": (6.33054e-01, 5), "slope": (3.685512e+05, 6), "charge": (-9.93246e-10, 5),
    "ratio": (3.169938e+00, 6),
}
MEASURE_POINTS = 1029

RC_LOWPASS_SP = """RC low-pass filter
* 1 kOhm with 159 nF: corner near 1 kHz
V1 in 0 DC 0 AC 1 PULSE(0 1 0 1u 1u 5m 10m)
R1 in out 1k
C1 out 0 159n
.op
.ac dec 10 10 100k
.tran 10u 5m
.end
"""
RC_LOWPASS_POINTS = 519

# docs/examples/diode_iv.sp, and the first rows of the DC plot the
# reference prints for it (docs/learn/analyses.md:86-93)
DIODE_IV_SP = """Diode I-V curve with a DC sweep
V1 a 0 DC 0
D1 a 0 dmod
.model dmod D(IS=1e-14 N=1)
.dc V1 0 0.8 0.01
.end
"""
DIODE_IV_ROWS = ["0\t0e0\t0e0\t0e0",
                 "1\t1e-2\t-1.386624089968383e-14\t1e-2",
                 "2\t2e-2\t-3.166792969038269e-14\t2e-2"]

CMOS_INVERTER_SP = """CMOS inverter
.param vdd=1.8 wn=1u wp={2*wn}
VDD vdd 0 {vdd}
VIN in 0 DC 0 PULSE(0 {vdd} 1n 100p 100p 5n 10n)
M1 out in 0 0 nch W={wn} L=180n
M2 out in vdd vdd pch W={wp} L=180n
CL out 0 10f
.model nch NMOS(LEVEL=1 VTO=0.45 KP=200u LAMBDA=0.05 GAMMA=0.4 PHI=0.8)
.model pch PMOS(LEVEL=1 VTO=-0.45 KP=80u LAMBDA=0.05 GAMMA=0.4 PHI=0.8)
.dc VIN 0 1.8 0.01
.tran 10p 20n
.meas tran tphl TRIG v(in) VAL=0.9 RISE=1 TARG v(out) VAL=0.9 FALL=1
.end
"""
CMOS_INVERTER_GOLDEN = {"tphl": (2.818778e-11, 6), "tphl.targ": (1.078188e-09, 6), "tphl.trig": (1.050000e-09, 6)}
CMOS_INVERTER_POINTS = 2032

# measure.sp with its R and C as .param, so points can move them
DIODE_RECTIFIER_SP = """Half-wave rectifier
* A 50 Hz, 10 V peak sine through a diode into an RC load.
V1 in 0 SIN(0 10 50)
D1 in out dmod
R1 out 0 1k
C1 out 0 100u
.model dmod D(IS=1e-14 N=1.05 RS=0.5 CJO=10p)
.tran 0.1m 60m
.meas tran vpeak MAX v(out)
.meas tran vmin MIN v(out) FROM=40m TO=60m
.end
"""

DIODE_RECTIFIER_GOLDEN = {"vpeak": (9.228482e+00, 6), "vpeak.at": (5.173860e-03, 6), "vmin": (7.719858e+00, 6),
                          "vmin.at": (4.317386e-02, 6)}
DIODE_RECTIFIER_POINTS = 610

SOURCES_SP = """Independent source waveforms
* Each source drives its own 1k load, so every node shows its waveform.
VP p 0 PULSE(0 1 1m 0.1m 0.1m 2m 5m)
VS s 0 SIN(0.5 0.5 1k)
VW w 0 PWL(0 0 1m 1 3m 1 4m 0)
VE e 0 EXP(0 1 1m 0.5m 3m 0.5m)
IC 0 c DC 1m
RP p 0 1k
RS s 0 1k
RW w 0 1k
RE e 0 1k
RC c 0 1k
.tran 10u 5m
.end
"""
SOURCES_POINTS = 526
# the same four waveforms as current sources into the same loads
ISOURCES_SP = """Independent current source waveforms
IP 0 p PULSE(0 1m 1m 0.1m 0.1m 2m 5m)
IS 0 s SIN(0.5m 0.5m 1k)
IW 0 w PWL(0 0 1m 1m 3m 1m 4m 0)
IE 0 e EXP(0 1m 1m 0.5m 3m 0.5m)
RP p 0 1k
RS s 0 1k
RW w 0 1k
RE e 0 1k
.tran 10u 5m
.end
"""  # docs/learn/sources.md:55-58: 10 devices, 10 variables

MONTE_AGAUSS_SP = """Divider spread with agauss and a Monte Carlo DC
* agauss(nominal, variation, sigmas): 30 ohms is the 3-sigma spread.
.param r1v=agauss(1k, 30, 3) r2v=agauss(1k, 30, 3)
V1 in 0 DC 1
R1 in out {r1v}
R2 out 0 {r2v}
.option seed=7
.dc monte=200
.end
"""

# docs/learn/monte-carlo.md:62-69 (`--rawfile --format=print`): the first
# three trials, (run, i(v1), v(in), v(out)), as the reference printed them.
MONTE_AGAUSS_ROWS = [
    (1, -4.98898735103262e-4, 1.0, 4.9498315376777796e-1),
    (2, -4.97205754614453e-4, 1.0, 4.945092576673322e-1),
    (3, -5.003152460377415e-4, 1.0, 4.957336298692346e-1),
]

MONTE_TRAN_SP = """RC delay spread over Monte Carlo trials
* R varies by 3 % (one sigma). The 50 % crossing is at R C ln 2 = 693 ns.
.param rv=agauss(1k, 30, 1)
V1 in 0 PWL(0 0 1p 1)
R1 in out 'rv'
C1 out 0 1n
.option seed=3
.tran 10n 2u sweep monte=50
.meas tran tc WHEN v(out)=0.5 RISE=1
.end
"""

# docs/learn/monte-carlo.md:80-90: the last two trials' tc lines and the
# statistics over the 50 trials, as the reference printed them.
MONTE_TRAN_TAIL = ["tc                  =   6.81503e-07", "tc                  =   6.87338e-07"]
MONTE_TRAN_STATS = ["mean(tc) = 6.894056e-07", "sigma(tc) = 2.106016e-08", "min(tc) = 6.229492e-07",
                    "max(tc) = 7.412147e-07"]

MEASURE_PARAM_SP = MEASURE_SP.replace("R1 in out 1k", ".param r=1k c=1n\nR1 in out {r}").replace(
    "C1 out 0 1n", "C1 out 0 {c}")
