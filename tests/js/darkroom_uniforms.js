'use strict'
// PtDevelop's darkroom words for the develop settings given as JSON in argv[2] (js/darkroom.js lightUniforms), as JSON.
const D = require('../../shadytracer_amd/js/darkroom.js')
const u = D.lightUniforms(JSON.parse(process.argv[2] || '{}'))
console.log(JSON.stringify({ ...u, curveLut: u.curveLut ? Array.from(u.curveLut) : null }))
