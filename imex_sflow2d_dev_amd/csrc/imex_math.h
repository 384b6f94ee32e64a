// Correctly rounded x**y for the FP64 path, in double-double arithmetic.
//
// The reference evaluates REAL**REAL through the C library's pow
// (Schiller-Naumann drag Re**0.687, src/constitutive_2d.f90:5378; hindered
// settling (1 - a/a_max)**4.65, :4850-4851; entrainment Ri**2.4, :5049).  A
// device pow with a 1-ulp error bound gives a different last bit in a
// fraction of the calls, and the settling iteration's 1e-6 stop and the
// Newton tolerances amplify such bits along a run.  pow_cr returns the
// correctly rounded result (round to nearest) of the exact x**y instead:
//   log x = e ln2 + log c + 2 atanh(s),  x = 2**e m,  m in [sqrt(1/2), sqrt(2)),
//           c = 1 + k/128 the nearest table point to m,  s = (m - c)/(m + c),
//           |s| < 0.0028, the atanh series summed to s**15;
//   exp z = 2**e 2**(j/128) (1 + expm1(r)),  z = (128 e + j) ln2/128 + r,
//           |r| < 0.0028, the expm1 Taylor series summed to r**11;
// with the tables log(1 + k/128) and 2**(j/128), and ln2 split in three
// parts, all exact to about 2**-106.  The relative error of the
// double-double result is below 2**-100, so the final rounding is correct
// unless x**y lies within 2**-47 ulp of a rounding boundary.  Every function
// here compiles for the host as well as the device (the host build is what
// the CPU tests check against a quad-precision pow).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace imex {
namespace ddm {

struct dd {
  double hi, lo;
};

__host__ __device__ __forceinline__ dd two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  return dd{s, e};
}

// requires |a| >= |b| (or a == 0)
__host__ __device__ __forceinline__ dd fast_two_sum(double a, double b) {
  const double s = a + b;
  return dd{s, b - (s - a)};
}

__host__ __device__ __forceinline__ dd two_prod(double a, double b) {
  const double p = a * b;
  return dd{p, fma(a, b, -p)};
}

__host__ __device__ __forceinline__ dd add(dd a, dd b) {
  dd s = two_sum(a.hi, b.hi);
  const dd t = two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = fast_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return fast_two_sum(s.hi, s.lo);
}

__host__ __device__ __forceinline__ dd mul(dd a, dd b) {
  dd p = two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return fast_two_sum(p.hi, p.lo);
}

__host__ __device__ __forceinline__ dd mul_d(dd a, double b) {
  dd p = two_prod(a.hi, b);
  p.lo += a.lo * b;
  return fast_two_sum(p.hi, p.lo);
}

__host__ __device__ __forceinline__ dd div(dd a, dd b) {
  const double q1 = a.hi / b.hi;
  dd r = add(a, mul_d(b, -q1));
  const double q2 = r.hi / b.hi;
  r = add(r, mul_d(b, -q2));
  const double q3 = r.hi / b.hi;
  return add(fast_two_sum(q1, q2), dd{q3, 0.0});
}

// 1/(2k+1), k = 0..7
__host__ __device__ __forceinline__ dd inv_odd(int k) {
  static constexpr double T[8][2] = {
      {0x1.0000000000000p+0, 0x0.0p+0},
      {0x1.5555555555555p-2, 0x1.5555555555555p-56},
      {0x1.999999999999ap-3, -0x1.999999999999ap-57},
      {0x1.2492492492492p-3, 0x1.2492492492492p-57},
      {0x1.c71c71c71c71cp-4, 0x1.c71c71c71c71cp-58},
      {0x1.745d1745d1746p-4, -0x1.745d1745d1746p-59},
      {0x1.3b13b13b13b14p-4, -0x1.3b13b13b13b14p-58},
      {0x1.1111111111111p-4, 0x1.1111111111111p-60}};
  return dd{T[k][0], T[k][1]};
}

// 1/n!, n = 2..11
__host__ __device__ __forceinline__ dd inv_fact(int n) {
  static constexpr double T[10][2] = {
      {0x1.0000000000000p-1, 0x0.0p+0},
      {0x1.5555555555555p-3, 0x1.5555555555555p-57},
      {0x1.5555555555555p-5, 0x1.5555555555555p-59},
      {0x1.1111111111111p-7, 0x1.1111111111111p-63},
      {0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65},
      {0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73},
      {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76},
      {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73},
      {0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76},
      {0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80}};
  return dd{T[n - 2][0], T[n - 2][1]};
}

// log(1 + k/128), k = -38..53
__host__ __device__ __forceinline__ dd log_table(int k) {
  static constexpr double T[92][2] = {
      {-0x1.68ac83e9c6a14p-2, -0x1.a64eadd740178p-58},
      {-0x1.5d5bddf595f30p-2, 0x1.6541148cbb8a2p-56},
      {-0x1.522ae0738a3d8p-2, 0x1.8f7e9b38a6979p-57},
      {-0x1.4718dc271c41bp-2, -0x1.8fb4c14c56eefp-60},
      {-0x1.3c25277333184p-2, 0x1.2ad27e50a8ec6p-56},
      {-0x1.314f1e1d35ce4p-2, 0x1.3d69909e5c3dcp-56},
      {-0x1.269621134db92p-2, -0x1.e0efadd9db02bp-56},
      {-0x1.1bf99635a6b95p-2, 0x1.12aeb84249223p-57},
      {-0x1.1178e8227e47cp-2, 0x1.0e63a5f01c691p-57},
      {-0x1.07138604d5862p-2, -0x1.cdb16ed4e9138p-56},
      {-0x1.f991c6cb3b379p-3, -0x1.f665066f980a2p-57},
      {-0x1.e530effe71012p-3, -0x1.2276041f43042p-59},
      {-0x1.d1037f2655e7bp-3, -0x1.60629242471a2p-57},
      {-0x1.bd087383bd8adp-3, -0x1.dd355f6a516d7p-60},
      {-0x1.a93ed3c8ad9e3p-3, -0x1.bcafa9de97203p-57},
      {-0x1.95a5adcf7017fp-3, -0x1.142c507fb7a3dp-58},
      {-0x1.823c16551a3c2p-3, 0x1.1232ce70be781p-57},
      {-0x1.6f0128b756abcp-3, 0x1.8de59c21e166cp-57},
      {-0x1.5bf406b543db2p-3, 0x1.1f5b44c0df7e7p-61},
      {-0x1.4913d8333b561p-3, 0x1.0d5604930f135p-58},
      {-0x1.365fcb0159016p-3, -0x1.7d411a5b944adp-58},
      {-0x1.23d712a49c202p-3, 0x1.6e38161051d69p-57},
      {-0x1.1178e8227e47cp-3, 0x1.0e63a5f01c691p-58},
      {-0x1.fe89139dbd566p-4, 0x1.ac9f4215f9393p-58},
      {-0x1.da727638446a2p-4, -0x1.401fa71733019p-58},
      {-0x1.b6ac88dad5b1cp-4, 0x1.0057eed1ca59fp-59},
      {-0x1.9335e5d594989p-4, 0x1.478a85704ccb7p-58},
      {-0x1.700d30aeac0e1p-4, 0x1.72566212cdd05p-61},
      {-0x1.4d3115d207eacp-4, -0x1.769f42c7842ccp-58},
      {-0x1.2aa04a44717a5p-4, 0x1.d15d38d2fa3f7p-58},
      {-0x1.08598b59e3a07p-4, 0x1.dd7009902bf32p-58},
      {-0x1.ccb73cdddb2ccp-5, 0x1.e48fb0500efd4p-59},
      {-0x1.894aa149fb343p-5, -0x1.a8be97660a23dp-60},
      {-0x1.466aed42de3eap-5, 0x1.cdd6f7f4a137ep-59},
      {-0x1.0415d89e74444p-5, -0x1.c05cf1d753622p-59},
      {-0x1.8492528c8cabfp-6, 0x1.d192d0619fa67p-60},
      {-0x1.0205658935847p-6, -0x1.27c8e8416e71fp-60},
      {-0x1.010157588de71p-7, -0x1.46662d417ced0p-62},
      {0x0.0p+0, 0x0.0p+0},
      {0x1.fe02a6b106789p-8, -0x1.e44b7e3711ebfp-67},
      {0x1.fc0a8b0fc03e4p-7, -0x1.83092c59642a1p-62},
      {0x1.7b91b07d5b11bp-6, -0x1.5b602ace3a510p-60},
      {0x1.f829b0e783300p-6, 0x1.33e3f04f1ef23p-60},
      {0x1.39e87b9febd60p-5, -0x1.5bfa937f551bbp-59},
      {0x1.77458f632dcfcp-5, 0x1.18d3ca87b9296p-59},
      {0x1.b42dd711971bfp-5, -0x1.eb9759c130499p-60},
      {0x1.f0a30c01162a6p-5, 0x1.85f325c5bbacdp-59},
      {0x1.16536eea37ae1p-4, -0x1.79da3e8c22cdap-60},
      {0x1.341d7961bd1d1p-4, -0x1.b599f227becbbp-58},
      {0x1.51b073f06183fp-4, 0x1.a49e39a1a8be4p-58},
      {0x1.6f0d28ae56b4cp-4, -0x1.906d99184b992p-58},
      {0x1.8c345d6319b21p-4, -0x1.4a697ab3424a9p-61},
      {0x1.a926d3a4ad563p-4, 0x1.942f48aa70ea9p-58},
      {0x1.c5e548f5bc743p-4, 0x1.5d617ef8161b1p-60},
      {0x1.e27076e2af2e6p-4, -0x1.61578001e0162p-60},
      {0x1.fec9131dbeabbp-4, -0x1.5746b9981b36cp-58},
      {0x1.0d77e7cd08e59p-3, 0x1.9a5dc5e9030acp-57},
      {0x1.1b72ad52f67a0p-3, 0x1.483023472cd74p-58},
      {0x1.29552f81ff523p-3, 0x1.301771c407dbfp-57},
      {0x1.371fc201e8f74p-3, 0x1.de6cb62af18a0p-58},
      {0x1.44d2b6ccb7d1ep-3, 0x1.9f4f6543e1f88p-57},
      {0x1.526e5e3a1b438p-3, -0x1.746ff8a470d3ap-57},
      {0x1.5ff3070a793d4p-3, -0x1.bc60efafc6f6ep-58},
      {0x1.6d60fe719d21dp-3, -0x1.caae268ecd179p-57},
      {0x1.7ab890210d909p-3, 0x1.be36b2d6a0608p-59},
      {0x1.87fa06520c911p-3, -0x1.bf7fdbfa08d9ap-57},
      {0x1.9525a9cf456b4p-3, 0x1.d904c1d4e2e26p-57},
      {0x1.a23bc1fe2b563p-3, 0x1.93711b07a998cp-59},
      {0x1.af3c94e80bff3p-3, -0x1.398cff3641985p-58},
      {0x1.bc286742d8cd6p-3, 0x1.4fce744870f55p-58},
      {0x1.c8ff7c79a9a22p-3, -0x1.4f689f8434012p-57},
      {0x1.d5c216b4fbb91p-3, 0x1.6e443597e4d40p-57},
      {0x1.e27076e2af2e6p-3, -0x1.61578001e0162p-59},
      {0x1.ef0adcbdc5936p-3, 0x1.48637950dc20dp-57},
      {0x1.fb9186d5e3e2bp-3, -0x1.caaae64f21acbp-57},
      {0x1.0402594b4d041p-2, -0x1.28ec217a5022dp-57},
      {0x1.0a324e27390e3p-2, 0x1.7dcfde8061c03p-56},
      {0x1.1058bf9ae4ad5p-2, 0x1.89fa0ab4cb31dp-58},
      {0x1.1675cababa60ep-2, 0x1.ce63eab883717p-61},
      {0x1.1c898c16999fbp-2, -0x1.0e5c62aff1c44p-60},
      {0x1.22941fbcf7966p-2, -0x1.76f5eb09628afp-56},
      {0x1.2895a13de86a3p-2, 0x1.7ad24c13f040ep-56},
      {0x1.2e8e2bae11d31p-2, -0x1.8f4cdb95ebdf9p-56},
      {0x1.347dd9a987d55p-2, -0x1.4dd4c580919f8p-57},
      {0x1.3a64c556945eap-2, -0x1.c68651945f97cp-57},
      {0x1.404308686a7e4p-2, -0x1.0bcfb6082ce6dp-56},
      {0x1.4618bc21c5ec2p-2, 0x1.f42decdeccf1dp-56},
      {0x1.4be5f957778a1p-2, -0x1.259b35b04813dp-57},
      {0x1.51aad872df82dp-2, 0x1.3927ac19f55e3p-59},
      {0x1.5767717455a6cp-2, 0x1.526adb283660cp-56},
      {0x1.5d1bdbf5809cap-2, 0x1.4236383dc7fe1p-56},
      {0x1.62c82f2b9c795p-2, 0x1.7b7af915300e5p-57},
  };
  return dd{T[k + 38][0], T[k + 38][1]};
}

// 2**(j/128), j = 0..127
__host__ __device__ __forceinline__ dd exp2_table(int j) {
  static constexpr double T[128][2] = {
      {0x1.0000000000000p+0, 0x0.0p+0},
      {0x1.0163da9fb3335p+0, 0x1.b61299ab8cdb7p-54},
      {0x1.02c9a3e778061p+0, -0x1.19083535b085dp-56},
      {0x1.04315e86e7f85p+0, -0x1.0a31c1977c96ep-54},
      {0x1.059b0d3158574p+0, 0x1.d73e2a475b465p-55},
      {0x1.0706b29ddf6dep+0, -0x1.c91dfe2b13c27p-55},
      {0x1.0874518759bc8p+0, 0x1.186be4bb284ffp-57},
      {0x1.09e3ecac6f383p+0, 0x1.1487818316136p-54},
      {0x1.0b5586cf9890fp+0, 0x1.8a62e4adc610bp-54},
      {0x1.0cc922b7247f7p+0, 0x1.01edc16e24f71p-54},
      {0x1.0e3ec32d3d1a2p+0, 0x1.03a1727c57b53p-59},
      {0x1.0fb66affed31bp+0, -0x1.b9bedc44ebd7bp-57},
      {0x1.11301d0125b51p+0, -0x1.6c51039449b3ap-54},
      {0x1.12abdc06c31ccp+0, -0x1.1b514b36ca5c7p-58},
      {0x1.1429aaea92de0p+0, -0x1.32fbf9af1369ep-54},
      {0x1.15a98c8a58e51p+0, 0x1.2406ab9eeab0ap-55},
      {0x1.172b83c7d517bp+0, -0x1.19041b9d78a76p-55},
      {0x1.18af9388c8deap+0, -0x1.11023d1970f6cp-54},
      {0x1.1a35beb6fcb75p+0, 0x1.e5b4c7b4968e4p-55},
      {0x1.1bbe084045cd4p+0, -0x1.95386352ef607p-54},
      {0x1.1d4873168b9aap+0, 0x1.e016e00a2643cp-54},
      {0x1.1ed5022fcd91dp+0, -0x1.1df98027bb78cp-54},
      {0x1.2063b88628cd6p+0, 0x1.dc775814a8495p-55},
      {0x1.21f49917ddc96p+0, 0x1.2a97e9494a5eep-55},
      {0x1.2387a6e756238p+0, 0x1.9b07eb6c70573p-54},
      {0x1.251ce4fb2a63fp+0, 0x1.ac155bef4f4a4p-55},
      {0x1.26b4565e27cddp+0, 0x1.2bd339940e9d9p-55},
      {0x1.284dfe1f56381p+0, -0x1.a4c3a8c3f0d7ep-54},
      {0x1.29e9df51fdee1p+0, 0x1.612e8afad1255p-55},
      {0x1.2b87fd0dad990p+0, -0x1.10adcd6381aa4p-59},
      {0x1.2d285a6e4030bp+0, 0x1.0024754db41d5p-54},
      {0x1.2ecafa93e2f56p+0, 0x1.1ca0f45d52383p-56},
      {0x1.306fe0a31b715p+0, 0x1.6f46ad23182e4p-55},
      {0x1.32170fc4cd831p+0, 0x1.a9ce78e18047cp-55},
      {0x1.33c08b26416ffp+0, 0x1.32721843659a6p-54},
      {0x1.356c55f929ff1p+0, -0x1.b5cee5c4e4628p-55},
      {0x1.371a7373aa9cbp+0, -0x1.63aeabf42eae2p-54},
      {0x1.38cae6d05d866p+0, -0x1.e958d3c9904bdp-54},
      {0x1.3a7db34e59ff7p+0, -0x1.5e436d661f5e3p-56},
      {0x1.3c32dc313a8e5p+0, -0x1.efff8375d29c3p-54},
      {0x1.3dea64c123422p+0, 0x1.ada0911f09ebcp-55},
      {0x1.3fa4504ac801cp+0, -0x1.7d023f956f9f3p-54},
      {0x1.4160a21f72e2ap+0, -0x1.ef3691c309278p-58},
      {0x1.431f5d950a897p+0, -0x1.1c7dde35f7999p-55},
      {0x1.44e086061892dp+0, 0x1.89b7a04ef80d0p-59},
      {0x1.46a41ed1d0057p+0, 0x1.c944bd1648a76p-54},
      {0x1.486a2b5c13cd0p+0, 0x1.3c1a3b69062f0p-56},
      {0x1.4a32af0d7d3dep+0, 0x1.9cb62f3d1be56p-54},
      {0x1.4bfdad5362a27p+0, 0x1.d4397afec42e2p-56},
      {0x1.4dcb299fddd0dp+0, 0x1.8ecdbbc6a7833p-54},
      {0x1.4f9b2769d2ca7p+0, -0x1.4b309d25957e3p-54},
      {0x1.516daa2cf6642p+0, -0x1.f768569bd93efp-55},
      {0x1.5342b569d4f82p+0, -0x1.07abe1db13cadp-55},
      {0x1.551a4ca5d920fp+0, -0x1.d689cefede59bp-55},
      {0x1.56f4736b527dap+0, 0x1.9bb2c011d93adp-54},
      {0x1.58d12d497c7fdp+0, 0x1.295e15b9a1de8p-55},
      {0x1.5ab07dd485429p+0, 0x1.6324c054647adp-54},
      {0x1.5c9268a5946b7p+0, 0x1.c4b1b816986a2p-60},
      {0x1.5e76f15ad2148p+0, 0x1.ba6f93080e65ep-54},
      {0x1.605e1b976dc09p+0, -0x1.3e2429b56de47p-54},
      {0x1.6247eb03a5585p+0, -0x1.383c17e40b497p-54},
      {0x1.6434634ccc320p+0, -0x1.c483c759d8933p-55},
      {0x1.6623882552225p+0, -0x1.bb60987591c34p-54},
      {0x1.68155d44ca973p+0, 0x1.038ae44f73e65p-57},
      {0x1.6a09e667f3bcdp+0, -0x1.bdd3413b26456p-54},
      {0x1.6c012750bdabfp+0, -0x1.2895667ff0b0dp-56},
      {0x1.6dfb23c651a2fp+0, -0x1.bbe3a683c88abp-57},
      {0x1.6ff7df9519484p+0, -0x1.83c0f25860ef6p-55},
      {0x1.71f75e8ec5f74p+0, -0x1.16e4786887a99p-55},
      {0x1.73f9a48a58174p+0, -0x1.0a8d96c65d53cp-54},
      {0x1.75feb564267c9p+0, -0x1.0245957316dd3p-54},
      {0x1.780694fde5d3fp+0, 0x1.866b80a02162dp-54},
      {0x1.7a11473eb0187p+0, -0x1.41577ee04992fp-55},
      {0x1.7c1ed0130c132p+0, 0x1.f124cd1164dd6p-54},
      {0x1.7e2f336cf4e62p+0, 0x1.05d02ba15797ep-56},
      {0x1.80427543e1a12p+0, -0x1.27c86626d972bp-54},
      {0x1.82589994cce13p+0, -0x1.d4c1dd41532d8p-54},
      {0x1.8471a4623c7adp+0, -0x1.8d684a341cdfbp-55},
      {0x1.868d99b4492edp+0, -0x1.fc6f89bd4f6bap-54},
      {0x1.88ac7d98a6699p+0, 0x1.994c2f37cb53ap-54},
      {0x1.8ace5422aa0dbp+0, 0x1.6e9f156864b27p-54},
      {0x1.8cf3216b5448cp+0, -0x1.0d55e32e9e3aap-56},
      {0x1.8f1ae99157736p+0, 0x1.5cc13a2e3976cp-55},
      {0x1.9145b0b91ffc6p+0, -0x1.dd6792e582524p-54},
      {0x1.93737b0cdc5e5p+0, -0x1.75fc781b57ebcp-57},
      {0x1.95a44cbc8520fp+0, -0x1.64b7c96a5f039p-56},
      {0x1.97d829fde4e50p+0, -0x1.d185b7c1b85d1p-54},
      {0x1.9a0f170ca07bap+0, -0x1.173bd91cee632p-54},
      {0x1.9c49182a3f090p+0, 0x1.c7c46b071f2bep-56},
      {0x1.9e86319e32323p+0, 0x1.824ca78e64c6ep-56},
      {0x1.a0c667b5de565p+0, -0x1.359495d1cd533p-54},
      {0x1.a309bec4a2d33p+0, 0x1.6305c7ddc36abp-54},
      {0x1.a5503b23e255dp+0, -0x1.d2f6edb8d41e1p-54},
      {0x1.a799e1330b358p+0, 0x1.bcb7ecac563c7p-54},
      {0x1.a9e6b5579fdbfp+0, 0x1.0fac90ef7fd31p-54},
      {0x1.ac36bbfd3f37ap+0, -0x1.f9234cae76cd0p-55},
      {0x1.ae89f995ad3adp+0, 0x1.7a1cd345dcc81p-54},
      {0x1.b0e07298db666p+0, -0x1.bdef54c80e425p-54},
      {0x1.b33a2b84f15fbp+0, -0x1.2805e3084d708p-57},
      {0x1.b59728de5593ap+0, -0x1.c71dfbbba6de3p-54},
      {0x1.b7f76f2fb5e47p+0, -0x1.5584f7e54ac3bp-56},
      {0x1.ba5b030a1064ap+0, -0x1.efcd30e54292ep-54},
      {0x1.bcc1e904bc1d2p+0, 0x1.23dd07a2d9e84p-55},
      {0x1.bf2c25bd71e09p+0, -0x1.efdca3f6b9c73p-54},
      {0x1.c199bdd85529cp+0, 0x1.11065895048ddp-55},
      {0x1.c40ab5fffd07ap+0, 0x1.b4537e083c60ap-54},
      {0x1.c67f12e57d14bp+0, 0x1.2884dff483cadp-54},
      {0x1.c8f6d9406e7b5p+0, 0x1.1acbc48805c44p-56},
      {0x1.cb720dcef9069p+0, 0x1.503cbd1e949dbp-56},
      {0x1.cdf0b555dc3fap+0, -0x1.dd83b53829d72p-55},
      {0x1.d072d4a07897cp+0, -0x1.cbc3743797a9cp-54},
      {0x1.d2f87080d89f2p+0, -0x1.d487b719d8578p-54},
      {0x1.d5818dcfba487p+0, 0x1.2ed02d75b3707p-55},
      {0x1.d80e316c98398p+0, -0x1.11ec18beddfe8p-54},
      {0x1.da9e603db3285p+0, 0x1.c2300696db532p-54},
      {0x1.dd321f301b460p+0, 0x1.2da5778f018c3p-54},
      {0x1.dfc97337b9b5fp+0, -0x1.1a5cd4f184b5cp-54},
      {0x1.e264614f5a129p+0, -0x1.7b627817a1496p-54},
      {0x1.e502ee78b3ff6p+0, 0x1.39e8980a9cc8fp-55},
      {0x1.e7a51fbc74c83p+0, 0x1.2d522ca0c8de2p-54},
      {0x1.ea4afa2a490dap+0, -0x1.e9c23179c2893p-54},
      {0x1.ecf482d8e67f1p+0, -0x1.c93f3b411ad8cp-54},
      {0x1.efa1bee615a27p+0, 0x1.dc7f486a4b6b0p-54},
      {0x1.f252b376bba97p+0, 0x1.3a1a5bf0d8e43p-54},
      {0x1.f50765b6e4540p+0, 0x1.9d3e12dd8a18bp-54},
      {0x1.f7bfdad9cbe14p+0, -0x1.dbb12d006350ap-54},
      {0x1.fa7c1819e90d8p+0, 0x1.74853f3a5931ep-55},
      {0x1.fd3c22b8f71f1p+0, 0x1.2eb74966579e7p-57},
  };
  return dd{T[j][0], T[j][1]};
}

// ln 2 = LN2_1 + LN2_2 + LN2_3 to about 2**-160
constexpr double LN2_1 = 0x1.62e42fefa39efp-1;
constexpr double LN2_2 = 0x1.abc9e3b39803fp-56;
constexpr double LN2_3 = 0x1.7b57a079a1934p-111;
constexpr double SQRT_HALF = 0x1.6a09e667f3bcdp-1;

// natural logarithm of a positive finite x
__host__ __device__ __forceinline__ dd log_dd(double x) {
  int e = 0;
  double m = frexp(x, &e);  // x = m 2**e, m in [0.5, 1)
  if (m < SQRT_HALF) {
    m = 2.0 * m;
    e -= 1;
  }
  const int k = (int)rint((m - 1.0) * 128.0);  // -38..53
  const double c = 1.0 + k * 0x1.0p-7;
  // m - c is exact (c/2 <= m <= 2c)
  const dd s = div(dd{m - c, 0.0}, two_sum(m, c));
  const dd s2 = mul(s, s);
  dd acc = inv_odd(7);
#pragma unroll 1
  for (int i = 6; i >= 0; i--) acc = add(inv_odd(i), mul(s2, acc));
  dd lm = mul(s, acc);
  lm.hi = 2.0 * lm.hi;
  lm.lo = 2.0 * lm.lo;
  lm = add(log_table(k), lm);
  const double ed = (double)e;
  dd el = add(two_prod(ed, LN2_1), two_prod(ed, LN2_2));
  el = add(el, dd{ed * LN2_3, 0.0});
  return add(el, lm);
}

// exp of a double-double argument, rounded to the nearest double
__host__ __device__ __forceinline__ double exp_dd(dd z) {
  if (z.hi > 709.8) return HUGE_VAL;
  if (z.hi < -745.2) return 0.0;
  const double nd = rint(z.hi * 0x1.71547652b82fep+7);  // z 128 / ln2
  const int n = (int)nd;
  const int j = n & 127;
  const int e = (n - j) / 128;
  dd r = add(z, two_prod(-nd, LN2_1 * 0x1.0p-7));
  r = add(r, two_prod(-nd, LN2_2 * 0x1.0p-7));
  r = add(r, dd{-nd * (LN2_3 * 0x1.0p-7), 0.0});
  dd acc = inv_fact(11);
#pragma unroll 1
  for (int i = 10; i >= 2; i--) acc = add(inv_fact(i), mul(r, acc));
  const dd p = add(r, mul(mul(r, r), acc));  // expm1(r)
  const dd t = exp2_table(j);
  const dd res = add(t, mul(t, p));
  if (e > -1022) return ldexp(res.hi, e);
  // result below 2**-1021: round the pair once on the grid of 2**-1074 --
  // scale so that grid is the integers (exact), round hi to an integer, then
  // let the remainder hi - f + lo decide the final step
  const double hs = ldexp(res.hi, e + 1074), ls = ldexp(res.lo, e + 1074);
  double f = rint(hs);
  const double rem = (hs - f) + ls;
  if (rem > 0.5 || (rem == 0.5 && fmod(f, 2.0) != 0.0)) f += 1.0;
  else if (rem < -0.5 || (rem == -0.5 && fmod(f, 2.0) != 0.0)) f -= 1.0;
  return ldexp(f, -1074);
}

// The double-double x**y above, kept out of line: pow_fast settles all but
// about one call in 10**8.
__host__ __device__ __noinline__ inline double pow_slow(double x, double y) {
  return exp_dd(mul_d(log_dd(x), y));
}

// ln2 = LN2H + LN2L (to 2**-102) with LN2H of 42 bits (e * LN2H is exact for
// |e| <= 2048); ln2/128 = LN2_128H + LN2_128L + LN2_128LL (to 2**-154) with
// LN2_128H of 36 bits (n * LN2_128H is exact for |n| < 2**17)
constexpr double LN2H = 0x1.62e42fefa3800p-1;
constexpr double LN2L = 0x1.ef35793c76730p-45;
constexpr double LN2_128H = 0x1.62e42fefa0000p-8;
constexpr double LN2_128L = 0x1.cf79abc9e3b3ap-47;
constexpr double LN2_128LL = -0x1.ff0342542fc33p-101;
// 2/3 and 1/6 as pairs (to 2**-108)
constexpr double C3H = 0x1.5555555555555p-1, C3L = 0x1.5555555555555p-55;
constexpr double C6H = 0x1.5555555555555p-3, C6L = 0x1.5555555555555p-57;

// 1/d to within 2**-51 relative for d in [1, 4]: the hardware estimate on
// the device (v_rcp_f64) and two Newton steps, which square its error
__host__ __device__ __forceinline__ double recip(double d) {
#ifdef __HIP_DEVICE_COMPILE__
  double r = __builtin_amdgcn_rcp(d);
#else
  double r = 1.0 / d;
#endif
  r = fma(r, fma(-d, r, 1.0), r);
  return fma(r, fma(-d, r, 1.0), r);
}

// The two tables as one array of POW_TAB_N pairs: log(1 + (i - 38)/128)
// for i < 92, then 2**((i - 92)/128).  A kernel that evaluates x**y often
// copies them to LDS (pow_tables_to_lds) and passes that copy to pow_cr:
// from global memory each lookup is a dependent load whose latency, when the
// line has left the vector L1, is several times that of LDS.
constexpr int POW_TAB_N = 92 + 128;
__host__ __device__ __forceinline__ dd pow_tab_entry(int i) {
  return (i < 92) ? log_table(i - 38) : exp2_table(i - 92);
}

// Ziv's strategy, first step: x**y as the pair R = R.hi + R.lo and a bound
// eps on its relative error, with the same tables and double arithmetic
// whose every rounding below 2**-85 of the result is made exact by
// error-free sums and products.  Returns false when the argument of exp
// leaves (-708, 708).
//   log x = e ln2 + log c + 2 atanh(s), s = (m - c)/(m + c) = sh + sl,
//           |s| < 2**-8.48:  2 atanh(s) = 2s + (2/3)s**3 + (2/5)s**5 +
//           (2/7)s**7 + (2/9)s**9 (truncation < 2**-88 |2s|), the s**3 term
//           and its sl part exact to 2**-100, the higher terms in double
//           (< 2**-87 |2s|); |2s| < 2.4 |log x| wherever the table point c
//           is not 1, so the logarithm is within 2**-85 |log x|;
//   z = y log x as a pair (error < 2**-85 |z| + 2**-104 |z|);
//   exp z = 2**e2 2**(j/128) exp(r), r = z - n ln2/128 = r1 + r2 exact to
//           2**-104 |z|, |r1| < 2**-8.47, |r2| < 2**-62:
//           exp(r) - 1 = r1 + r1**2/2 + r1**3 (1/6 + r1 P(r1)) + r2 (1 + r1
//           + r1**2/2), P to r1**5/40320 (truncation < 2**-94), the first
//           three terms exact to 2**-89;
// so eps = 2**-84 |z| + 2**-87 holds with a margin of two.
__host__ __device__ __forceinline__ bool pow_fast_pair(double x, double y, dd& R, double& eps,
                                                       int& e2, const dd* tab = nullptr) {
  int e = 0;
  double m = frexp(x, &e);
  if (m < SQRT_HALF) {
    m = 2.0 * m;
    e -= 1;
  }
  const int k = (int)rint((m - 1.0) * 128.0);
  const double c = 1.0 + k * 0x1.0p-7;
  const double num = m - c;      // exact
  const dd den = two_sum(m, c);  // den.hi in [1.4, 2.9]
  const double rden = recip(den.hi);
  const double sh = num * rden;
  const double sl = (fma(-sh, den.hi, num) - sh * den.lo) * rden;
  const dd t = two_prod(sh, sh);
  const dd s3 = two_prod(sh, t.hi);
  const dd p3 = two_prod(C3H, s3.hi);
  const double p3l = p3.lo + (C3H * (s3.lo + sh * t.lo) + C3L * s3.hi);
  const double p5 = (s3.hi * t.hi) *
                    (0x1.999999999999ap-2 + t.hi * (0x1.2492492492492p-2 + t.hi * 0x1.c71c71c71c71cp-3));
  const dd lc = tab ? tab[k + 38] : log_table(k);
  const double ed = (double)e;
  const dd u = two_sum(ed * LN2H, lc.hi);
  const dd v = two_sum(u.hi, 2.0 * sh);
  const dd w = two_sum(v.hi, p3.hi);
  const double lo =
      u.lo + v.lo + w.lo + (ed * LN2L + lc.lo + (2.0 * sl + (p3l + (2.0 * t.hi * sl + p5))));
  const dd l = fast_two_sum(w.hi, lo);
  dd z = two_prod(y, l.hi);
  z = fast_two_sum(z.hi, z.lo + y * l.lo);
  if (!(fabs(z.hi) < 708.0)) return false;
  const double nd = rint(z.hi * 0x1.71547652b82fep+7);  // z 128 / ln2, |nd| < 2**17
  const int n = (int)nd;
  const int j = n & 127;
  e2 = (n - j) / 128;
  // r = z - nd ln2/128: the first difference is exact (Sterbenz), the
  // product with LN2_128L error-free
  const double a = z.hi - nd * LN2_128H;
  const dd b = two_prod(-nd, LN2_128L);
  const dd r0 = two_sum(a, b.hi);
  const dd r = two_sum(r0.hi, r0.lo + (z.lo + (b.lo - nd * LN2_128LL)));
  const double r1 = r.hi, r2 = r.lo;
  const dd q2 = two_prod(r1, r1);
  const double q2h = 0.5 * q2.hi, q2l = 0.5 * q2.lo;
  const dd c3 = two_prod(r1, q2.hi);
  const double c3l = c3.lo + r1 * q2.lo;
  const double P = 0x1.5555555555555p-5 +
                   r1 * (0x1.1111111111111p-7 +
                         r1 * (0x1.6c16c16c16c17p-10 + r1 * (0x1.a01a01a01a01ap-13 + r1 * 0x1.a01a01a01a01ap-16)));
  const dd p = two_prod(c3.hi, C6H);
  const double pl = p.lo + (c3.hi * C6L + c3l * C6H + c3.hi * (r1 * P));
  const double small = q2l + pl + r2 * (1.0 + r1 + q2h);
  const dd T = tab ? tab[92 + j] : exp2_table(j);
  const dd mp = two_prod(T.hi, r1);
  const dd mq = two_prod(T.hi, q2h);
  const dd mpp = two_prod(T.hi, p.hi);
  const dd a1 = two_sum(T.hi, mp.hi);
  const dd a2 = two_sum(mq.hi, mpp.hi);
  const dd a3 = two_sum(a1.hi, a2.hi);
  const double rl = a1.lo + a3.lo + a2.lo +
                    (mp.lo + mq.lo + mpp.lo + (T.hi * small + (T.lo + T.lo * (r1 + (q2h + p.hi)))));
  R = fast_two_sum(a3.hi, rl);
  eps = fabs(z.hi) * 0x1.0p-84 + 0x1.0p-87;
  return true;
}

// x**y when R rounds the same way at both ends of its error interval;
// otherwise false (the caller then takes pow_slow)
__host__ __device__ __forceinline__ bool pow_fast(double x, double y, double& out,
                                                  const dd* tab = nullptr) {
  dd R;
  double eps;
  int e2;
  if (!pow_fast_pair(x, y, R, eps, e2, tab)) return false;
  int ex = 0;
  const double fr = frexp(R.hi, &ex);  // R.hi = fr 2**ex, |fr| in [0.5, 1)
  double half_ulp = ldexp(1.0, ex - 54);
  if (fabs(fr) == 0.5) half_ulp *= 0.5;  // a power of two: the gap below is halved
  // the factor covers the rounding of the left-hand side
  if (!(fabs(R.lo) + eps * fabs(R.hi) < (1.0 - 0x1.0p-50) * half_ulp)) return false;
  if (ex + e2 < -1020 || ex + e2 > 1023) return false;  // leave the range ends to pow_slow
  out = ldexp(R.hi, e2);
  return true;
}

// Correctly rounded log, log1p and exp from the double-double routines above
// (their pairs are within 2**-100 of the exact values, so the rounding of
// the pair is the correct one unless the exact value lies within about
// 2**-47 ulp of a rounding boundary).  Used by the pore-pressure terms,
// where the reference reaches the C library's cpow.
__host__ __device__ inline double log_cr(double x) { return log_dd(x).hi; }
__host__ __device__ inline double exp_cr(double x) { return exp_dd(dd{x, 0.0}); }
// log(1 + u) for u > -1: 1 + u = hi + lo exactly, log(hi) + log1p(lo/hi)
// with lo/hi as a pair and log1p(e) = e - e**2/2 (|e| < 2**-53)
__host__ __device__ inline double log1p_cr(double u) {
  const dd w = two_sum(1.0, u);
  const dd L = log_dd(w.hi);
  const dd e = div(dd{w.lo, 0.0}, dd{w.hi, 0.0});
  const dd t = add(e, dd{-0.5 * (e.hi * e.hi), 0.0});
  return add(L, t).hi;
}
// x*x + y*y - 1 rounded once (the pair sum is exact to 2**-104)
__host__ __device__ inline double x2y2m1_cr(double x, double y) {
  return add(add(two_prod(x, x), dd{-1.0, 0.0}), two_prod(y, y)).hi;
}

}  // namespace ddm

// x**y correctly rounded; the special cases follow C's pow.  tab: the tables
// in LDS (pow_tables_to_lds) or null for the ones in global memory.
__host__ __device__ __forceinline__ double pow_cr(double x, double y,
                                                  const ddm::dd* tab = nullptr) {
  if (y == 0.0 || x == 1.0) return 1.0;
  if (x == 0.0) return (y > 0.0) ? 0.0 : HUGE_VAL;
  if (!(x > 0.0) || !(x < HUGE_VAL) || !(fabs(y) < HUGE_VAL)) return pow(x, y);
  double r;
  if (ddm::pow_fast(x, y, r, tab)) return r;
  return ddm::pow_slow(x, y);
}

// the block's copy of the x**y tables in LDS; every thread of the block
// calls this (it ends with a barrier)
__device__ __forceinline__ void pow_tables_to_lds(ddm::dd* t) {
  for (int i = (int)threadIdx.x; i < ddm::POW_TAB_N; i += (int)blockDim.x) t[i] = ddm::pow_tab_entry(i);
  __syncthreads();
}

}  // namespace imex
