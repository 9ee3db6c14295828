'use strict'
// The optics bench through the N-API addon: scenes as JSON on argv[2], results as JSON on stdout, for
// tests/test_optics_gpu.py.
const path = require('path')
const fs = require('fs')
const { createOpticsBench } = require(path.join(__dirname, '..', '..', 'shadytracer_amd', 'js', 'optics'))

const scenes = JSON.parse(fs.readFileSync(process.argv[2], 'utf8'))
const O = createOpticsBench({ device: 0 })
const { results, far, walkMs } = O.run(scenes, { far: true })
let farSum = 0
for (let i = 0; i < far.length; i++) farSum += far[i]
process.stdout.write(JSON.stringify({ results, farSum, walkMs }))
