#!/usr/bin/env node
'use strict';
/**
 * IRT chemistry on a saved rad_buf dump, on the GPU: the counterpart of the reference's
 * tools/run_irt.cjs (usage and output contract of run_irt.cjs:1-79).
 *
 *   node run_irt.js <dump_file> <n_therm> <E_eV> [seed]
 *
 * dump_file: raw little-endian float32 (x, y, z, pid*8+code) per radical, as POSTed to the
 * reference's dump server (tools/dump_server.cjs:38-54).  Prints one JSON line
 * {type: 'result', timeline, n_reacted, chem_n, t_wall, n_repaired, rxn_info} on stdout and
 * the summary of run_irt.cjs:40-51 on stderr.  Exit 1: usage; 2: no GPU device.
 */
const fs = require('fs');
const host = require('./dna_host');

async function main(argv) {
  const dumpFile = argv[2];
  const nTherm = parseInt(argv[3] || '0', 10);
  const E = parseInt(argv[4] || '10000', 10);
  const seed = parseInt(argv[5] || '1', 10);
  if (!dumpFile || !nTherm) {
    console.error('usage: node run_irt.js <dump_file> <n_therm> <E_eV> [seed]');
    return 1;
  }
  const buf = fs.readFileSync(dumpFile);
  const rad = new Float32Array(buf.buffer.slice(buf.byteOffset, buf.byteOffset + buf.length));
  const radN = rad.length / 4;
  console.error(`[run_irt] ${dumpFile}: ${radN} radicals (${(buf.length / 1e6).toFixed(2)} MB)  n_therm=${nTherm}  E=${E} eV`);
  if (!host.ensurePipelines(0)) {
    console.error('[run_irt] no GPU device');
    return 2;
  }
  const res = await host.runChemistryWorker(rad, radN, nTherm, E, undefined, undefined, seed);
  const out = {
    type: 'result', timeline: res.timeline, n_reacted: res.n_reacted, chem_n: res.chem_n, t_wall: res.t_wall,
    n_repaired: 0, rxn_info: res.rxn_info,
  };
  process.stdout.write(JSON.stringify(out) + '\n');
  console.error(`[worker] ${res.n_reacted} reactions in ${(res.t_wall / 1000).toFixed(1)}s`);
  for (const cp of res.timeline) {
    console.error(`  ${cp.label.padEnd(8)} G(OH)=${cp.G_OH.toFixed(3).padStart(7)}  G(eaq)=${cp.G_eaq.toFixed(3).padStart(7)}  ` +
      `G(H)=${cp.G_H.toFixed(3).padStart(7)}  G(H2O2)=${cp.G_H2O2.toFixed(3).padStart(7)}  G(H2)=${cp.G_H2.toFixed(3).padStart(7)}`);
  }
  console.error('  reactions fired:');
  for (const rx of res.rxn_info) {
    console.error(`    ${rx.label.padEnd(28)} count=${String(rx.count).padStart(7)}  σ=${rx.sigma}nm  rc=${rx.rc}nm`);
  }
  return 0;
}

main(process.argv).then((rc) => process.exit(rc), (e) => { console.error(e); process.exit(1); });
