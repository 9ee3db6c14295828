'use strict'
/*
 * Command-line front of the emitter: prints one JSON object { scene, camera, settings } for a named
 * configuration (configs.js) or a builder document file.
 *
 *   node emit.js <config-name> [json-options]
 *   node emit.js --doc <scene.json> [camera-json]
 */
const fs = require('fs')
const { generate } = require('./hipgen')
const { CONFIGS } = require('./configs')

function main(argv) {
    let made
    if (argv[0] === '--doc') {
        const doc = JSON.parse(fs.readFileSync(argv[1], 'utf8'))
        made = { doc, camera: argv[2] ? JSON.parse(argv[2]) : null, settings: {} }
    } else {
        const f = CONFIGS[argv[0]]
        if (!f) throw new Error(`unknown configuration '${argv[0]}' (have ${Object.keys(CONFIGS).join(', ')})`)
        made = f(argv[1] ? JSON.parse(argv[1]) : {})
    }
    const scene = generate(made.doc)
    scene.params = Array.from(scene.params)
    process.stdout.write(JSON.stringify({ scene, camera: made.camera, settings: made.settings, doc: made.doc }))
}

try { main(process.argv.slice(2)) } catch (e) { process.stderr.write(String(e && e.stack || e) + '\n'); process.exit(1) }
