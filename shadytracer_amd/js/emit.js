'use strict'
/*
 * Command-line front of the emitter: prints one JSON object { scene, camera, settings } for a named
 * configuration (configs.js) or a builder document file.
 *
 *   node emit.js <config-name> [json-options]
 *   node emit.js --doc <scene.json> [camera-json | json-options]
 *
 * A document's options: camera (the engine's pinhole; without one, the document's own view), width, height
 * (256 x 256), approximate (render without what the engine does not draw, listed in scene.approximations).
 */
const fs = require('fs')
const { generate, cameraOf } = require('./hipgen')
const { CONFIGS } = require('./configs')

function main(argv) {
    let made, opts = {}
    if (argv[0] === '--doc') {
        const doc = JSON.parse(fs.readFileSync(argv[1], 'utf8'))
        const a = argv[2] ? JSON.parse(argv[2]) : {}
        opts = a && a.pos ? { camera: a } : a
        const w = opts.width || 256, h = opts.height || 256
        const skip = []
        const camera = opts.camera || (doc.view ? cameraOf(doc, w, h, (where, what) => skip.push(`${where}: ${what}`)) : null)
        if (skip.length && !opts.approximate) throw new Error(`hipgen: the engine does not draw ${skip.join('; ')}`)
        made = { doc, camera, settings: { width: w, height: h } }
    } else {
        const f = CONFIGS[argv[0]]
        if (!f) throw new Error(`unknown configuration '${argv[0]}' (have ${Object.keys(CONFIGS).join(', ')})`)
        made = f(argv[1] ? JSON.parse(argv[1]) : {})
    }
    const scene = generate(made.doc, { approximate: !!opts.approximate })
    scene.params = Array.from(scene.params)
    process.stdout.write(JSON.stringify({ scene, camera: made.camera, settings: made.settings, doc: made.doc }))
}

try { main(process.argv.slice(2)) } catch (e) { process.stderr.write(String(e && e.stack || e) + '\n'); process.exit(1) }
