'use strict';
/**
 * run_validation.js — command-line runValidation (src/app.ts:91-161) on the MI355X
 * library: prints one JSON object {rows, results} (results in the layout of
 * validation/webgpu-results.json) and logs to stderr.
 *
 *   node run_validation.js [--np N] [--energies 1000,10000] [--box NM] [--chem worker|none]
 *                          [--indirect-ssb]   (score the OH alive at 1 us too)
 */
const host = require('./dna_host');

function arg(name, dflt) {
  const i = process.argv.indexOf(name);
  return i >= 0 && i + 1 < process.argv.length ? process.argv[i + 1] : dflt;
}

async function main() {
  const np = parseInt(arg('--np', '4096'), 10);
  const energies = arg('--energies', null);
  const cfg = {
    np, boxNm: parseFloat(arg('--box', '15000')), ceEV: host.SOLVATION_THRESHOLD,
    chemBackend: arg('--chem', 'worker'),
    energies: energies ? energies.split(',').map(Number) : undefined,
    indirectSSB: process.argv.includes('--indirect-ssb'),
    log: (m, cls) => process.stderr.write(`[${cls || ''}] ${m}\n`),
  };
  const out = await host.runValidation(cfg);
  if (out === null) {
    process.stderr.write('no GPU device\n');
    process.exitCode = 3;
    return;
  }
  process.stdout.write(JSON.stringify(out) + '\n');
}

main().catch((e) => { process.stderr.write(`${e.stack || e}\n`); process.exitCode = 1; });
