// Physical, namelist and WSM6 constants of the ARW path (values of
// pywrf/config_params.py:29-253).  Kept as double so that products of two
// constants fold in double precision the way the reference's Python scalars
// do; kernels round each folded value once to float with F().
#pragma once

#define F(x) ((float)(x))

namespace pw {
// namelist (config_params.py:29-49, 114-118)
constexpr double spec_bdy_width = 5, spec_zone = 1, relax_zone = 4;
constexpr double smdiv = 0.1, emdiv = 0.01, kvdif = 0., khdif = 0.;
constexpr double dampcoef = 0.2, zdamp = 5000.;
constexpr int time_step = 60;
// species slots (config_params.py:101-112)
constexpr int P_QV = 1, P_QC = 2, P_QR = 3, P_QI = 4, P_QS = 5, P_QG = 6, n_moist = 7, n_scalar = 3;
// physics (config_params.py:120-162)
constexpr double r_d = 287.0, p1000mb = 100000.0, t00 = 300.0, p00 = 100000.0, p0 = 100000.0, t0 = 300.0;
constexpr double g = 9.810000, rcp = 0.2857143;
constexpr double SVP1 = 0.611199975013732910, SVP2 = 17.670000076293945312, SVP3 = 29.649999618530273438;
constexpr double SVPT0 = 273.149993896484375, t0c = 273.149993896484375, EP_2 = 0.621750414371490479;
constexpr double XLV = 2500000., prandtl = 1. / 3., c_s = 0.2500000, Cp = 1004.500, r_v = 461.6000;
constexpr double cpovcv = 1.40000, reradius = 1.5698588e-7, epssm = 0.1, rvovrd = 1.608362, Cpv = 1846.400;
constexpr double cvpm = -0.7142857;   // config_params.py:124 (-cv / cp, as written there)
constexpr double ep_1 = 0.608362436294555664, ep_2 = 0.621750414371490479, epsilon = 1.0000000e-15;
constexpr double XLS = 2850000., XLF = 350000.0, rhoair0 = 1.280000, rhowater = 1000.000;
constexpr double cliq = 4190.000, cice = 2106.000, psat = 610.7800;
// WSM6 (config_params.py:164-253)
constexpr double dtcldcr = 120., n0r = 8.0e6, avtr = 841.9, bvtr = 0.8, r0 = 0.8e-5, peaut = .55;
constexpr double xncr = 3.0e8, xmyu = 1.718e-5, avts = 11.72, bvts = 0.41, n0smax = 1.0e11;
constexpr double lamdarmax = 8.0e4, lamdasmax = 1.0e5, dicon = 11.9, dimax = 500.0e-6, n0s = 2.0e6;
constexpr double alpha = 0.12, pfrz1 = 100., pfrz2 = 0.66, qcrmin = 1.0e-9, eacrc = 1.0, dens = 100.0;
constexpr double qs0 = 6.0e-4, qc0 = 5.0265482E-04, qck1 = 6.773895, pidnc = 523.5988;
constexpr double bvtr1 = 1.800000, bvtr2 = 2.900000, bvtr3 = 3.800000, bvtr4 = 4.800000;
constexpr double g1pbr = 0.9312252, g3pbr = 4.690783, g4pbr = 17.81730, g5pbro2 = 1.826591;
constexpr double pvtr = 2500.064, eacrr = 1.000000, pacrr = 2.4813369E+10, bvtr6 = 6.800000;
constexpr double g6pbr = 495.4590, precr1 = 3.9207076E+07, precr2 = 8.2585357E+08, roqimax = 8.1250037E-05;
constexpr double bvts1 = 1.410000, bvts2 = 2.705000, bvts3 = 3.410000, bvts4 = 4.410000;
constexpr double g1pbs = 0.8866756, g3pbs = 3.011565, g4pbs = 10.26533, n0g = 4000000.0;
constexpr double avtg = 330.0000, bvtg = 0.8000000, deng = 500.0000, lamdagmax = 60000.00;
constexpr double g5pbso2 = 1.550299, pvts = 20.05161, pacrs = 5.5442112E+07, precs1 = 5200000.0;
constexpr double precs2 = 1.8681944E+07, pidn0r = 2.5132743E+10, pidn0s = 6.2831853E+08, xlv1 = 2343.600;
constexpr double pacrc = 5.5442112E+07, pi = 3.141593;
constexpr double bvtg1 = 1.800000, bvtg2 = 2.900000, bvtg3 = 3.800000, bvtg4 = 4.800000;
constexpr double g1pbg = 0.9312252, g3pbg = 4.690783, g4pbg = 17.81730, g5pbgo2 = 1.826591;
constexpr double pvtg = 979.9514, pacrg = 4.8630548E+09, precg1 = 1.9603538E+07, precg2 = 2.5852333E+08;
constexpr double pidn0g = 6.2831857E+09;
constexpr double rslopermax = 1.2500000E-05, rslopesmax = 9.9999997E-06, rslopegmax = 1.6666667E-05;
constexpr double rsloperbmax = 1.1954404E-04, rslopesbmax = 8.9125093E-03, rslopegbmax = 1.5048006E-04;
constexpr double rsloper2max = 1.5624999E-10, rslopes2max = 9.9999994E-11, rslopeg2max = 2.7777777E-10;
constexpr double rsloper3max = 1.9531248E-15, rslopes3max = 9.9999990E-16, rslopeg3max = 4.6296297E-15;
}  // namespace pw
