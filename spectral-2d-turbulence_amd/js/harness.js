'use strict';
// GPU parity harness over the JS surface: the checks of the reference's browser harness
// (tests/index.html:114-207) on the same fixtures and tolerances, run through Solver / Fft2D.
// Usage: node harness.js [fixtures.json]   (exit status 0 when every check passes)
const fs = require('fs');
const path = require('path');
const { Solver, Fft2D } = require('./index.js');

function relError(got, want) {
  let num = 0;
  let den = 0;
  for (let i = 0; i < want.length; i++) {
    num = Math.max(num, Math.abs(got[i] - want[i]));
    den = Math.max(den, Math.abs(want[i]));
  }
  return den > 0 ? num / den : num;
}

async function main() {
  const file = process.argv[2] || path.join(__dirname, '..', '..', 'tests', 'golden', 'fixtures.json');
  const fx = JSON.parse(fs.readFileSync(file, 'utf8'));
  const n = fx.fft.n;
  let failures = 0;
  const report = (name, err, tol) => {
    const ok = err <= tol;
    if (!ok) failures += 1;
    console.log(`${ok ? 'PASS' : 'FAIL'}  ${name.padEnd(44)} ${err.toExponential(3)}  tol ${tol.toExponential(1)}`);
  };

  const solver = await Solver.create(null, { n, dyeScale: 1 });
  solver.configure({
    beta: fx.step.beta, drag: fx.step.mu, viscosity: 1 / 1.5, hyperP: fx.step.hyperP,
    forceAmp: 0, forceK: 20, forceWidth: 2, cfl: 0.5, dtMin: 1e-6, dtMax: 1, dyeFade: 1,
    fixedDt: fx.step.dt,
  });
  report('hyperviscosity coefficient matches reference', Math.abs(solver.nuH / fx.step.nuH - 1), 1e-6);

  const fft = await Fft2D.create(null, n);
  report('fft2 forward', relError(fft.forward(fx.fft.input), fx.fft.forward), 2e-5);
  report('fft2 inverse', relError(fft.inverse(fx.fft.input), fx.fft.inverse), 2e-5);

  const [a, b] = solver.velocityPack(fx.velocity.what);
  report('velocity packing (omega + i u)', relError(a, fx.velocity.packA), 2e-5);
  report('velocity packing (v)', relError(b, fx.velocity.packB), 2e-5);

  report('nonlinear assembly with dealiasing', relError(solver.assemble(fx.assemble.z), fx.assemble.rhs), 5e-5);
  solver.writeState(fx.velocity.what);
  report('fused rhs (row + column kernels)', relError(solver.rhs(), fx.assemble.rhs), 5e-5);

  solver.writeState(fx.step.before);
  solver.step(1);
  report('one integrating-factor RK3 step', relError(solver.readState(), fx.step.after), 5e-5);
  solver.writeState(fx.step.before);
  solver.step(20);
  report('twenty steps accumulated', relError(solver.readState(), fx.run20.after), 5e-3);
  const d = await solver.readDiagnostics();
  report('energy after twenty steps', Math.abs(d.energy / fx.run20.energy - 1), 1e-4);
  report('enstrophy after twenty steps', Math.abs(d.enstrophy / fx.run20.enstrophy - 1), 1e-4);

  solver.destroy();
  console.log(failures === 0 ? 'ALL CHECKS PASS' : `${failures} CHECK(S) FAILED`);
  process.exit(failures === 0 ? 0 : 1);
}

main().catch((err) => {
  console.error('ERROR:', err.message);
  process.exit(2);
});
